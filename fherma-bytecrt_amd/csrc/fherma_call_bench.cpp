// fherma_call_bench: times fherma_run at the reference harness's boundary
// (steady_clock around the call, native-batched/main.cpp:182-187): host
// std::vector operands in, owned std::vector result out.
//
//   fherma_call_bench B N W seed reps
//
// q: W random bits with the top and bottom bits set; operands uniform in
// [0, q) by rejection, from a splitmix64 stream keyed by seed.  Prints one
// JSON object: init seconds, every call's milliseconds, the median, and
// whether products 0 and B-1 equal a separate B=1 run of the same product.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fherma.h"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};

// Little-endian limb compare: x < q.
bool below(const uint32_t* x, const std::vector<uint32_t>& q) {
    for (size_t i = q.size(); i-- > 0;)
        if (x[i] != q[i]) return x[i] < q[i];
    return false;
}

void fill_operand(uint32_t* dst, size_t N, const std::vector<uint32_t>& q, uint32_t W, SplitMix& rng) {
    const size_t L = q.size();
    const uint32_t top_bits = W - 32 * uint32_t(L - 1);
    const uint32_t top_mask = top_bits == 32 ? 0xFFFFFFFFu : ((1u << top_bits) - 1);
    for (size_t i = 0; i < N; ++i) {
        uint32_t* x = dst + i * L;
        do {
            for (size_t l = 0; l < L; ++l) x[l] = uint32_t(rng.next());
            x[L - 1] &= top_mask;
        } while (!below(x, q));
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: fherma_call_bench B N W seed reps\n");
        return 2;
    }
    const uint32_t B = uint32_t(std::atoi(argv[1])), N = uint32_t(std::atoi(argv[2])), W = uint32_t(std::atoi(argv[3]));
    const uint64_t seed = std::strtoull(argv[4], nullptr, 10);
    const int reps = std::max(1, std::atoi(argv[5]));
    const uint32_t L = (W + 31) / 32;
    SplitMix rng{seed * 0x100000001B3ull + 17};

    fherma::Point point{};
    point.B = B;
    point.N = N;
    point.W = W;
    point.L = L;
    point.q.shape = {int64_t(L)};
    point.q.data.resize(L);
    for (auto& w : point.q.data) w = uint32_t(rng.next());
    const uint32_t top = (W - 1) % 32;
    point.q.data[L - 1] &= top == 31 ? 0xFFFFFFFFu : ((2u << top) - 1);
    point.q.data[L - 1] |= 1u << top;
    point.q.data[0] |= 1u;

    fherma::Inputs in;
    const size_t words = size_t(N) * L;
    for (auto* t : {&in.a, &in.b}) {
        t->shape = {int64_t(B), int64_t(N), int64_t(L)};
        t->data.resize(size_t(B) * words);
        for (uint32_t p = 0; p < B; ++p) fill_operand(t->data.data() + p * words, N, point.q.data, W, rng);
    }

    using Clock = std::chrono::steady_clock;
    try {
        const auto t0 = Clock::now();
        void* state = fherma_init(point);
        const double init_s = std::chrono::duration<double>(Clock::now() - t0).count();
        std::vector<double> ms;
        fherma::Outputs first;
        for (int r = 0; r < reps; ++r) {
            const auto s = Clock::now();
            fherma::Outputs out = fherma_run(state, in);
            ms.push_back(std::chrono::duration<double, std::milli>(Clock::now() - s).count());
            if (r == 0) first = std::move(out);
        }
        fherma_free(state);

        // Products 0 and B-1 again, each alone through a B=1 state.
        bool exact = first.c.data.size() == size_t(B) * words;
        fherma::Point one = point;
        one.B = 1;
        void* st1 = fherma_init(one);
        for (uint32_t p : {0u, B - 1}) {
            fherma::Inputs in1;
            in1.a.shape = in1.b.shape = {1, int64_t(N), int64_t(L)};
            in1.a.data.assign(in.a.data.begin() + p * words, in.a.data.begin() + (p + 1) * words);
            in1.b.data.assign(in.b.data.begin() + p * words, in.b.data.begin() + (p + 1) * words);
            const fherma::Outputs o = fherma_run(st1, in1);
            exact = exact && std::equal(o.c.data.begin(), o.c.data.end(), first.c.data.begin() + p * words);
        }
        fherma_free(st1);

        std::vector<double> sorted = ms;
        std::sort(sorted.begin(), sorted.end());
        std::printf("{\"B\": %u, \"N\": %u, \"W\": %u, \"init_s\": %.6f, \"median_ms\": %.4f, \"min_ms\": %.4f, "
                    "\"exact\": %s, \"ms\": [",
                    B, N, W, init_s, sorted[sorted.size() / 2], sorted.front(), exact ? "true" : "false");
        for (size_t i = 0; i < ms.size(); ++i) std::printf("%s%.4f", i ? ", " : "", ms[i]);
        std::printf("]}\n");
        return exact ? 0 : 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "fherma_call_bench: %s\n", e.what());
        return 1;
    }
}
