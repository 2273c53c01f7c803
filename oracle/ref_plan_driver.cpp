// Driver for the reference's host plan (native-smoke/rns-plan.hpp, compiled in
// place from the reference checkout).  Test infrastructure only.
//
//   ref_plan N W gap  ->  writes a little-endian binary dump to stdout:
//   u32 pcount, kpad, l; u32 primes[P], psi[P], omega[P], omega_inv[P],
//   n_inverse[P], crt_inverse[P]; f64 invp[P]; u8 b_matrix[7L*kpad];
//   u8 b_byte_matrix[4L*kpad].
#include "rns-plan.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: ref_plan N W gap\n");
        return 2;
    }
    const unsigned N = std::strtoul(argv[1], nullptr, 10);
    const unsigned W = std::strtoul(argv[2], nullptr, 10);
    const uint32_t gap = static_cast<uint32_t>(std::strtoul(argv[3], nullptr, 10));
    const unsigned L = (W + 31) / 32;
    std::vector<uint32_t> q(L, 0xffffffffu);
    q[0] = uint32_t{0} - gap;
    q[L - 1] = (uint32_t{1} << (W - 32 * (L - 1))) - 1;
    try {
        const bytecrt::RnsPlan plan = bytecrt::make_rns_plan(N, W, q);
        auto put = [](const void* p, size_t n) { std::fwrite(p, 1, n, stdout); };
        const uint32_t head[3] = {plan.pcount, plan.kpad, plan.l};
        put(head, sizeof(head));
        const size_t P = plan.pcount;
        put(plan.primes.data(), 4 * P);
        put(plan.psi_roots.data(), 4 * P);
        put(plan.omega_roots.data(), 4 * P);
        put(plan.omega_inverse.data(), 4 * P);
        put(plan.n_inverse.data(), 4 * P);
        put(plan.crt_inverse.data(), 4 * P);
        put(plan.invp.data(), 8 * P);
        put(plan.b_matrix.data(), plan.b_matrix.size());
        put(plan.b_byte_matrix.data(), plan.b_byte_matrix.size());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
