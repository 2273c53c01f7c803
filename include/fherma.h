// fherma.h -- the C++ boundary the reference's harness links against
// (native-batched/fherma.h:13-49, "polymul/negacyclic-batched@1.0.0").  The
// type and function declarations must stay identical for a drop-in; they are
// implemented on top of the C ABI in include/bytecrt.h
// (fherma-bytecrt_amd/csrc/fherma_solve.cpp).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace fherma {

template <class T>
struct Tensor {
    std::vector<int64_t> shape;
    std::vector<T> data;  // row-major

    int64_t count() const {
        int64_t total = 1;
        for (auto d : shape) total *= d;
        return total;
    }
};

struct Point {
    uint32_t B = 0;         // products per call
    uint32_t N = 0;         // coefficients per polynomial
    uint32_t W = 0;         // modulus bits
    uint32_t L = 0;         // u32 limbs per coefficient
    Tensor<uint32_t> q;     // tensor<L x u32>
};

struct Inputs {
    Tensor<uint32_t> a;     // tensor<B x N x L x u32>
    Tensor<uint32_t> b;     // tensor<B x N x L x u32>
};

struct Outputs {
    Tensor<uint32_t> c;     // tensor<B x N x L x u32>
};

}  // namespace fherma

void* fherma_init(const fherma::Point& p);
fherma::Outputs fherma_run(void* state, const fherma::Inputs& in);
void fherma_free(void* state);
