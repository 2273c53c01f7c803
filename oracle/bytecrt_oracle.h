/*
 * bytecrt_oracle.h -- CPU restatement of the ByteCRT product (test
 * infrastructure only: tests, the smoke entry and the bench CPU baseline load
 * it as the checker; the product library never links it).
 */
#ifndef BYTECRT_ORACLE_H
#define BYTECRT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_P 128
#define ORC_MAX_L 64

typedef struct OrcPlan OrcPlan;

int orc_prime32(uint32_t p);
OrcPlan* orc_plan_create(uint32_t N, uint32_t W, const uint32_t* q, uint32_t L, char* err, int errlen);
void orc_plan_free(OrcPlan* pl);
uint32_t orc_plan_pcount(const OrcPlan* pl);
uint32_t orc_plan_kpad(const OrcPlan* pl);
uint32_t orc_plan_product_bits(const OrcPlan* pl);
void orc_plan_tables(const OrcPlan* pl, uint32_t* primes, uint32_t* psi, uint32_t* omega,
                     uint32_t* omega_inv, uint32_t* n_inv, uint32_t* crt_inv);
void orc_plan_invp(const OrcPlan* pl, double* out);
void orc_plan_weights(const OrcPlan* pl, uint32_t* out);
void orc_plan_b7(const OrcPlan* pl, uint8_t* out);
void orc_plan_b4(const OrcPlan* pl, uint8_t* out);

void orc_encode(const OrcPlan* pl, const uint32_t* x, size_t count, uint32_t* out);
uint64_t orc_encode_bytepower(const OrcPlan* pl, const uint32_t* x, size_t count, uint32_t* out);

void orc_ntt_negacyclic(uint32_t p, uint32_t psi, uint32_t n, const uint32_t* a,
                        const uint32_t* b, uint32_t* c);
void orc_dense_input(uint32_t n, uint32_t p, uint32_t operand, uint32_t* a);

uint32_t orc_alpha_f32(const OrcPlan* pl, const uint32_t* terms);
uint32_t orc_alpha_f64(const OrcPlan* pl, const uint32_t* terms);
uint32_t orc_crt_b7_wide(const OrcPlan* pl, const uint32_t* terms_alpha, uint32_t* z);
void orc_crt_b4_wide(const OrcPlan* pl, const uint32_t* terms_alpha, uint32_t* z);
void orc_reduce_generic(const OrcPlan* pl, const uint32_t* z, uint32_t* r);
int orc_reduce_near_power(const OrcPlan* pl, const uint32_t* z, uint32_t* r);
void orc_crt_terms(const OrcPlan* pl, const uint32_t* terms, size_t count, uint32_t* out);

int orc_product(const OrcPlan* pl, const uint32_t* a, const uint32_t* b, uint32_t* c);

int orc_threads(void);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
