/*
 * dg_sgns_oracle.h - CPU oracle for the SGNS trainer.  TEST INFRASTRUCTURE
 * ONLY: loaded by the tests, the smoke check and the benchmark's CPU-baseline
 * leg, to check results, never by the product library.
 */
#ifndef DG_SGNS_ORACLE_H
#define DG_SGNS_ORACLE_H

#include <stdint.h>

int dgo_vocab(uint32_t v_num, const uint64_t* cn, uint32_t min_count, uint32_t* vertex_of_rank,
              uint32_t* rank_of_vertex, uint64_t* cn_of_rank, uint32_t* vocab_size,
              uint64_t* train_words);
void dgo_init_syn0(uint32_t vocab, uint32_t dim, int float_ops, float* syn0);
void dgo_unigram_table(const uint64_t* cn_of_rank, uint32_t vocab, uint32_t table_size,
                       int32_t* table);
void dgo_exp_table(float* t);
void dgo_keep_limits(const uint64_t* cn_of_rank, uint32_t vocab, float sample, uint64_t train_words,
                     uint32_t* keep);
uint64_t dgo_subsample(const uint32_t* tokens, const uint64_t* offsets, uint64_t n_walks,
                       const uint32_t* rank_of_vertex, const uint32_t* keep, uint64_t seed,
                       uint32_t epoch, int32_t* sen, uint64_t* sen_off);
void dgo_sg_plain_batch(const int32_t* sen, const uint64_t* sen_off, uint64_t n_sen,
                        uint32_t window, uint32_t negative, const int32_t* table,
                        uint32_t table_size, uint32_t vocab, float alpha, const float* expt,
                        uint32_t dim, float* syn0, float* syn1);
void dgo_sg_reuse_batch(const int32_t* sen, const uint64_t* sen_off, uint64_t n_sen,
                        uint32_t window, uint32_t negative, const int32_t* negs, float alpha,
                        const float* expt, uint32_t dim, float* syn0, float* syn1);
void dgo_reuse_negs(uint64_t n_sen, uint32_t negative, uint64_t seed, uint32_t epoch,
                    const int32_t* table, uint32_t table_size, uint32_t vocab, int32_t* out);
uint64_t dgo_plain_batches(const uint64_t* sen_off, uint64_t n_sen, uint64_t* batch_first,
                           uint64_t cap);
/* Philox draws of one position: window shrink b and `draws` negatives. */
void dgo_position_draws(uint64_t seed, uint64_t gs, uint32_t q, uint32_t epoch, uint32_t window,
                        uint32_t draws, const int32_t* table, uint32_t table_size, uint32_t vocab,
                        uint32_t* b, int32_t* negs);
/* Batched form: sequential SGNS with per-position shared negatives, FP64. */
void dgo_sg_shared(const int32_t* sen, const uint64_t* sen_off, uint64_t first_sen, uint64_t n_sen,
                   uint32_t window, uint32_t negative, uint32_t draws, const int32_t* table,
                   uint32_t table_size, uint32_t vocab, float alpha, const float* expt,
                   int exact_sigmoid, uint64_t seed, uint32_t epoch, uint32_t dim, float* syn0,
                   float* syn1);
/* The batched forms' evaluation of that order (first order in the updates). */
void dgo_sg_shared_first_order(const int32_t* sen, const uint64_t* sen_off, uint64_t first_sen,
                               uint64_t n_sen, uint32_t window, uint32_t negative, uint32_t draws,
                               const int32_t* table, uint32_t table_size, uint32_t vocab,
                               float alpha, const float* expt, int exact_sigmoid, uint64_t seed,
                               uint32_t epoch, uint32_t dim, float* syn0, float* syn1);
/* Tuple interface, FP64: logits and loss, gradient rows, central
 * differences, and the reference's sequential in-place update. */
double dgo_tuple_forward(const uint32_t* tuples, uint64_t n, uint32_t K, const float* syn0,
                         const float* syn1, uint32_t dim, int bf16, double* logits);
void dgo_tuple_grad_rows(const uint32_t* tuples, uint64_t n, uint32_t K, const float* syn0,
                         const float* syn1, uint32_t dim, uint32_t vocab, int bf16,
                         const double* dlogits, int table, const uint32_t* rows, uint64_t n_rows,
                         double* out);
double dgo_tuple_fd(const uint32_t* tuples, uint64_t n, uint32_t K, const float* syn0,
                    const float* syn1, uint32_t dim, int table, uint32_t row, uint32_t col,
                    double h);
void dgo_tuple_inplace(const uint32_t* tuples, uint64_t n, uint32_t K, double lr, uint32_t dim,
                       double* syn0, double* syn1);

/* DSGL (include/dsgl.hpp:783-1427), see dsgl_oracle.c for what is kept and
 * what is chosen. */
typedef struct {
  const uint32_t* tokens;          /* corpus (vertex ids) */
  const uint64_t* offsets;         /* n_walks + 1 */
  uint64_t n_walks;
  const uint32_t* rank_of_vertex;  /* vocabulary rank of each vertex, or 0xffffffff */
  const uint32_t* keep;            /* per rank: dgo_dsgl_keep_limits */
  const int32_t* table;            /* unigram table */
  uint32_t table_size;
  uint32_t vocab;
  const float* expt;               /* sigmoid table (1001 entries) */
  int exact_sigmoid;
  uint32_t dim, window, negative, iter;
  uint32_t workers;                /* compute threads the corpus is split over */
  float alpha0;
  uint64_t train_words;            /* sum of the vocabulary's counts */
  uint64_t seed;
  uint32_t epoch0;                 /* epoch number of the first epoch (Philox key) */
} dgo_dsgl_args;
void dgo_dsgl_keep_limits(const uint64_t* cn_of_rank, uint32_t vocab, float sample,
                          uint64_t train_words, uint32_t* keep);
void dgo_dsgl_unigram_table(const uint64_t* cn_of_rank, uint32_t vocab, uint32_t table_size,
                            int32_t* table);
int dgo_dsgl_train_f64(const dgo_dsgl_args* a, float* syn0, float* syn1);
int dgo_dsgl_train_f32(const dgo_dsgl_args* a, float* syn0, float* syn1, int threads);

#endif
