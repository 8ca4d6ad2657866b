/*
 * ffo.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain CPU restatement of
 * fastfermion's truncated Heisenberg-picture propagation.  It exists to check
 * the HIP product path (tests, smoke check) and to serve as the timed CPU
 * baseline in bench.py.  Nothing in the product (fastfermion_amd/) links,
 * loads or calls it.
 *
 * Gate and option records have exactly the layout of the product's C ABI
 * records (include/fastfermion_amd.h) so that one packed gate table can be
 * fed to both sides of a parity test.
 *
 * Bit layout of a Pauli term with n_words = W (W in 1..4):
 *   x[i*W + w] : bit b set  <=>  qubit 64*w+b carries X or Y   (reference "xory")
 *   z[i*W + w] : bit b set  <=>  qubit 64*w+b carries Y or Z   (reference "yorz")
 * (src/pauli/algebra.h:30-33).  Majorana terms use one occupation mask m
 * with the same word order (src/majorana_algebra.h:30-36).
 */
#ifndef FFO_ORACLE_H
#define FFO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FFO_ROT = 0, FFO_H = 1, FFO_S = 2, FFO_CNOT = 3, FFO_SWAP = 4, FFO_CZ = 5 };
enum { FFO_SERIAL = 0, FFO_SHARDED = 1, FFO_OMP = 2 };

#define FFO_NO_DEGREE_LIMIT 0x7FFFFFFF

typedef struct ffo_pauli_gate {
    int32_t kind;      /* FFO_ROT or one of the Clifford kinds */
    int32_t q0, q1;    /* Clifford qubits (q1 unused for H/S) */
    int32_t reserved;
    double theta;      /* ROT angle: U = exp(-i theta/2 P) */
    uint64_t x[4];     /* ROT axis P, xory words */
    uint64_t z[4];     /* ROT axis P, yorz words */
} ffo_pauli_gate;

typedef struct ffo_majorana_gate {
    uint64_t m[4];     /* axis Majorana string (occupation words) */
    double theta;      /* U = exp(-i theta/2 * i^r m_S) */
    int32_t r;         /* 1 when m_S is not Hermitian (degree mod 4 in {2,3}) */
    int32_t reserved;
} ffo_majorana_gate;

typedef struct ffo_options {
    int32_t maxdegree;      /* partner degree filter, compared literally (a negative value drops
                               every partner); FFO_NO_DEGREE_LIMIT: no filter */
    int32_t topk;           /* keep the topk largest |c| per truncation (0 = off) */
    int32_t max_xweight;    /* X/Y-weight cap (-1 = off) */
    int32_t xtrunc_period;  /* apply the X/Y-weight cap every this many rotations */
    int32_t batched;        /* 1: truncate once per commuting run of rotations */
    int32_t precision;      /* ignored by the oracle (always FP64 complex) */
    double mincoeff;        /* erase |c| <= mincoeff when > 0 */
} ffo_options;

/* Result buffers, allocated by the oracle with malloc, released by ffo_free. */
typedef struct ffo_terms {
    int64_t n_terms;
    int32_t n_words;
    int32_t reserved;
    uint64_t* x;     /* term-major [n_terms][n_words]; Majorana: occupation mask */
    uint64_t* z;     /* term-major; NULL for Majorana */
    double* c_re;
    double* c_im;
    double seconds;      /* wall time of the propagation proper (steady clock) */
    int64_t updates;     /* term-gate updates: sum over rotations of live terms entering it */
    int64_t peak_terms;  /* largest live-term count seen after any rotation */
} ffo_terms;

/* Pauli propagation (src/pauli/propagate.h:297-493).  strategy: FFO_SERIAL,
 * FFO_SHARDED (key-hash sharded maps, n_threads shards) or FFO_OMP. */
int ffo_pauli_propagate(const ffo_pauli_gate* gates, int64_t n_gates, int32_t n_words,
                        const uint64_t* x, const uint64_t* z, const double* c_re,
                        const double* c_im, int64_t n_terms, const ffo_options* opts,
                        int32_t strategy, int32_t n_threads, ffo_terms* out);

/* Majorana propagation (src/majorana_propagation.h:18-55); maxdegree as in ffo_options. */
int ffo_majorana_propagate(const ffo_majorana_gate* gates, int64_t n_gates, int32_t n_words,
                           const uint64_t* m, const double* c_re, const double* c_im,
                           int64_t n_terms, int32_t maxdegree, double mincoeff, ffo_terms* out);

/* <b|O|b> for a qubit product state (src/qubitproductstate.h:73-88); b has n_words words. */
void ffo_pauli_expect(const ffo_terms* t, const uint64_t* b, double out[2]);

/* <b|O|b> for a Fock product state over modes (src/fockstate.h:95-114). */
void ffo_majorana_expect(const ffo_terms* t, const uint64_t* occ, double out[2]);

void ffo_free(ffo_terms* t);

/* Live terms entering each rotation of the calling thread's last propagate call,
 * in application order (serial strategy); returns the rotation count, writes at
 * most cap.  Propagate calls from different threads may run concurrently. */
int64_t ffo_rotation_counts(int64_t* out, int64_t cap);

/* Threads the OpenMP runtime would use by default (0 when built without OpenMP). */
int32_t ffo_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif
