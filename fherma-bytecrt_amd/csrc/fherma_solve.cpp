// fherma_init / fherma_run / fherma_free over the C ABI (include/bytecrt.h).
//
// Error behaviour follows the reference adapter: B outside 1..128 and wrong
// input sizes raise std::runtime_error (native-batched/solve.cu:161-162,
// 207-208); geometry and modulus problems raise std::invalid_argument
// (rns-plan.hpp:182-196); device failures raise std::runtime_error
// (core/solve.cu:46-49).  All messages start with "ByteCRT".
#include "../../include/fherma.h"

#include <memory>

#include "../../include/bytecrt.h"

namespace {

struct Solver {
    bcrt_engine* engine = nullptr;
    uint32_t B = 0, N = 0, L = 0;
    ~Solver() { bcrt_destroy(engine); }
};

[[noreturn]] void raise(int code, const std::string& msg) {
    if (code == BCRT_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

}  // namespace

void* fherma_init(const fherma::Point& p) {
    if (p.B == 0 || p.B > 128) throw std::runtime_error("ByteCRT batch coverage B=1..128");
    if (p.q.data.size() != p.L) throw std::invalid_argument("ByteCRT RNS modulus limb count");
    auto s = std::make_unique<Solver>();
    char err[256] = {0};
    const int rc = bcrt_create(&s->engine, p.B, p.N, p.W, p.L, p.q.data.data(), err, sizeof(err));
    if (rc != BCRT_OK) raise(rc, err);
    s->B = p.B;
    s->N = p.N;
    s->L = p.L;
    return s.release();
}

fherma::Outputs fherma_run(void* state, const fherma::Inputs& in) {
    auto& s = *static_cast<Solver*>(state);
    const size_t total = size_t(s.B) * s.N * s.L;
    if (in.a.data.size() != total || in.b.data.size() != total)
        throw std::runtime_error("ByteCRT input coefficient count");
    fherma::Outputs out;
    out.c.shape = {int64_t(s.B), int64_t(s.N), int64_t(s.L)};
    out.c.data.resize(total);
    const int rc = bcrt_run(s.engine, in.a.data.data(), in.b.data.data(), out.c.data.data(), s.B);
    if (rc != BCRT_OK) raise(rc, bcrt_last_error(s.engine));
    return out;
}

void fherma_free(void* state) { delete static_cast<Solver*>(state); }
