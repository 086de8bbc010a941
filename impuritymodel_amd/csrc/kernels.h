// Launchers of the engine's HIP kernels (declared here, defined in *.hip).
#pragma once

#include "common.h"

namespace imped {

// Compiled operator tables of several models (device pointers).
//   model m owns groups [model_group_off[m], model_group_off[m+1]);
//   group g owns terms [group_term_off[g], group_term_off[g+1]) and maps a
//   state s to s ^ group_flip[g];
//   term t: care/val/sgn masks at term_masks[t*3W + {0,W,2W}], a term is
//   non-zero on s iff (s & care) == val, its sign is
//   (-1)^(popcount(s & sgn) + (term_meta[t] & 1)); term_meta >> 1 is the
//   term's position in the operator dict; term_coef[t] its value.
//   model m checks restrictions [model_restr_off[m], model_restr_off[m+1]):
//   lo <= popcount(s & restr_mask[r]) <= hi.
struct OpsView {
  const int32_t* model_group_off;
  const int32_t* group_term_off;
  const uint64_t* group_flip;
  const uint64_t* term_masks;
  const uint32_t* term_meta;
  const double2* term_coef;
  const int32_t* model_restr_off;
  const uint64_t* restr_mask;
  const int32_t* restr_lohi;
};

// One wave's slice of source states, all of one model and one chain.
struct WaveTask {
  int32_t model;
  int32_t chain;
  int64_t begin;
  int64_t end;
};

// Open-addressing state -> id table over (chain, key); slots hold pool ids,
// candidate ids (| CAND_FLAG) or EMPTY.
constexpr int32_t HASH_EMPTY = -1;
constexpr int32_t CAND_FLAG = 0x40000000;

struct HashView {
  int32_t* table;
  uint32_t mask;
  const uint64_t* pool_keys;
  const int32_t* pool_chain;
};

struct CandView {
  uint64_t* keys;
  int32_t* chain;
  int32_t* slot;
  int32_t* count;
  int32_t* alive_list;
  int32_t* alive_count;
  int32_t cap;
};

enum ApplyMode { MODE_EXPAND = 0, MODE_FIXED = 1, MODE_VECTOR = 2 };

struct ApplyArgs {
  OpsView ops;
  const WaveTask* tasks;
  int64_t n_tasks;
  const uint64_t* src_keys;  // [n_src * W]
  const double2* src_amps;   // vector mode only
  const int32_t* src_pos;    // vector mode: position of the source in its vector
  double weight_min;
  int32_t* counts;           // count pass output, one per source
  const int64_t* offsets;    // fill pass input, one per source (exclusive scan)
  // expand fill
  uint64_t* out_keys;
  double2* out_val;
  int32_t* out_chain;
  // fixed fill
  int32_t* out_col;
  const int32_t* pool_local;
  // vector fill
  int32_t* out_vec;
  int32_t* out_src;
  int32_t* out_first;
  HashView hash;
  CandView cand;
};

void launch_apply(int W, int mode, bool fill, const ApplyArgs& a, hipStream_t s);
void launch_discover(int W, const uint64_t* keys, const int32_t* chain, int64_t E, HashView h, CandView c,
                     hipStream_t s);

void launch_hash_insert_pool(int W, HashView h, int64_t first, int64_t count, hipStream_t s);
void launch_hash_clear(int32_t* table, uint32_t cap, hipStream_t s);

// Build per-chain wave tasks for sources [begin, end) sorted by chain.
// chain_of_src[i] gives the chain of source begin+i; chain_model maps chains
// to models.  Returns the task count (synchronizes).
int64_t build_wave_tasks(const int32_t* chain_of_src, int64_t begin, int64_t end, const int32_t* chain_model,
                         int n_chains, void* scratch_tasks, size_t scratch_cap_tasks, hipStream_t s,
                         int32_t* d_tmp_heads);

// Exclusive scan of int32 counts into int64 offsets; returns the total.
int64_t scan_counts(const int32_t* counts, int64_t n, int64_t* offsets, hipStream_t s);

// Sort `n` unique (chain, key) records given by index list `perm` (in/out);
// keys at keys[idx*W], chains at chain[idx].
void sort_chain_keys(int W, int n_orb, int n_chain_bits, const uint64_t* keys, const int32_t* chain, int32_t* perm,
                     int64_t n, hipStream_t s);
// Stable sort of records (vec, key) by LSD passes; same interface.
void sort_vec_keys_stable(int W, int n_orb, int n_vec_bits, const uint64_t* keys, const int32_t* vec,
                          int32_t* perm, int64_t n, hipStream_t s);
// Stable sort of perm by keys[perm[i]] on bits [begin_bit, end_bit).
void sort_u64_perm(const uint64_t* keys, int32_t* perm, int64_t n, int begin_bit, int end_bit, hipStream_t s);
// Stable sort of perm by 64-bit secondary key then 32-bit primary key.
void sort_pairs_u64_u32(const uint64_t* sec, const int32_t* prim, int n_prim_bits, int32_t* perm, int64_t n,
                        hipStream_t s);

}  // namespace imped
