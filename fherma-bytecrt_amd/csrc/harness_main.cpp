// bytecrt_solution: runs a point directory through fherma_init / fherma_run /
// fherma_free with the reference harness's wire format
// (native-batched/main.cpp:3-15,106-133):
//
//   <dir>/manifest.json          {"B":..,"N":..,"W":..,"L":..,"cases":..}
//   <dir>/point/q.bin            L little-endian u32
//   <dir>/cases/NNNNNN/a.bin     B*N*L u32, b.bin likewise
//   <dir>/out/NNNNNN/c.bin       written
//   <dir>/out/results.json       {"init_s":..,"cases":[{"i":..,"seconds":..,"status":"ok"}, ...]}
//
// Only the fherma_run call is timed; results.json is rewritten after every
// case so a killed run keeps what it finished.  A case that fails to read,
// run or write is recorded as "crashed" with a note.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fherma.h"

namespace fs = std::filesystem;

namespace {

std::string read_file(const fs::path& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot read " + path.string());
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

// Value of "key": <number> in a flat JSON object.
double manifest_number(const std::string& json, const std::string& key) {
    const std::string quoted = "\"" + key + "\"";
    const auto at = json.find(quoted);
    if (at == std::string::npos) throw std::runtime_error("manifest.json has no " + key);
    const auto colon = json.find(':', at + quoted.size());
    if (colon == std::string::npos) throw std::runtime_error("manifest.json malformed at " + key);
    return std::strtod(json.c_str() + colon + 1, nullptr);
}

std::string case_name(size_t i) {
    char buf[16];
    std::snprintf(buf, sizeof(buf), "%06zu", i);
    return buf;
}

fherma::Tensor<uint32_t> read_tensor(const fs::path& path, std::vector<int64_t> shape) {
    fherma::Tensor<uint32_t> t;
    t.shape = std::move(shape);
    const std::string raw = read_file(path);
    const size_t count = size_t(t.count());
    if (raw.size() != count * sizeof(uint32_t))
        throw std::runtime_error(path.filename().string() + ": " + std::to_string(raw.size()) +
                                 " bytes for " + std::to_string(count) + " values");
    t.data.resize(count);
    std::memcpy(t.data.data(), raw.data(), raw.size());
    return t;
}

void write_tensor(const fs::path& path, const fherma::Tensor<uint32_t>& t) {
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char*>(t.data.data()), std::streamsize(t.data.size() * sizeof(uint32_t)));
    if (!out) throw std::runtime_error("cannot write " + path.string());
}

std::string json_escape(const std::string& s) {
    std::string o;
    for (char c : s.substr(0, 200)) {
        if (c == '"' || c == '\\') o += '\\';
        o += c == '\n' ? ' ' : c;
    }
    return o;
}

void write_results(const fs::path& out_dir, double init_s, const std::vector<std::string>& cases) {
    std::ofstream f(out_dir / "results.json");
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.9f", init_s);
    f << "{\"init_s\":" << buf << ",\"cases\":[";
    for (size_t i = 0; i < cases.size(); ++i) f << (i ? "," : "") << cases[i];
    f << "]}";
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: bytecrt_solution <point directory>" << std::endl;
        return 2;
    }
    const fs::path root = argv[1];
    fherma::Point point{};
    size_t total = 0;
    try {
        const std::string manifest = read_file(root / "manifest.json");
        point.B = uint32_t(manifest_number(manifest, "B"));
        point.N = uint32_t(manifest_number(manifest, "N"));
        point.W = uint32_t(manifest_number(manifest, "W"));
        point.L = uint32_t(manifest_number(manifest, "L"));
        total = size_t(manifest_number(manifest, "cases"));
        point.q = read_tensor(root / "point" / "q.bin", {int64_t(point.L)});
    } catch (const std::exception& e) {
        std::cerr << "bytecrt_solution: " << e.what() << std::endl;
        return 2;
    }
    const fs::path out_root = root / "out";
    fs::create_directories(out_root);

    const auto t0 = std::chrono::steady_clock::now();
    void* state = nullptr;
    try {
        state = fherma_init(point);
    } catch (const std::exception& e) {
        std::cerr << "bytecrt_solution: fherma_init: " << e.what() << std::endl;
        return 1;
    }
    const double init_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<std::string> cases;
    write_results(out_root, init_s, cases);

    const std::vector<int64_t> shape = {int64_t(point.B), int64_t(point.N), int64_t(point.L)};
    for (size_t i = 0; i < total; ++i) {
        const fs::path in_dir = root / "cases" / case_name(i);
        const fs::path ans_dir = out_root / case_name(i);
        auto crashed = [&](const std::string& why) {
            cases.push_back("{\"i\":" + std::to_string(i) + ",\"seconds\":null,\"status\":\"crashed\",\"note\":\"" +
                            json_escape(why) + "\"}");
            write_results(out_root, init_s, cases);
        };
        fherma::Inputs in;
        try {
            in.a = read_tensor(in_dir / "a.bin", shape);
            in.b = read_tensor(in_dir / "b.bin", shape);
        } catch (const std::exception& e) {
            crashed(std::string("reading the case: ") + e.what());
            continue;
        }
        fherma::Outputs out;
        double seconds = 0;
        try {
            const auto s = std::chrono::steady_clock::now();
            out = fherma_run(state, in);
            seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - s).count();
        } catch (const std::exception& e) {
            crashed(e.what());
            continue;
        }
        try {
            fs::create_directories(ans_dir);
            write_tensor(ans_dir / "c.bin", out.c);
        } catch (const std::exception& e) {
            crashed(std::string("writing the answer: ") + e.what());
            continue;
        }
        char buf[64];
        std::snprintf(buf, sizeof(buf), "%.9f", seconds);
        cases.push_back("{\"i\":" + std::to_string(i) + ",\"seconds\":" + buf + ",\"status\":\"ok\"}");
        write_results(out_root, init_s, cases);
    }
    fherma_free(state);
    write_results(out_root, init_s, cases);
    return 0;
}
