// cbm_internal.h -- structures shared by the step kernels and the host side
// of libcbm_hip.so.  Not part of the C ABI.
#ifndef CBM_INTERNAL_H
#define CBM_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/cbm_hip.h"
#include "cbm_structs.h"

namespace cbm {

// Per-agent columns moved by the cell sort permutation.
struct PermCols {
  void* cur[MAX_COLS];
  void* alt[MAX_COLS];
  int32_t bytes[MAX_COLS];
  int32_t n;
};


// ------------------------------------------------------ kernel arguments
struct RebuildArgs {
  RowGeom g;
  const int32_t* flag;        // NULL: unconditional
  int32_t* flag_clear;        // flag of the next step, cleared by k_keys
  int64_t* rebuilds;          // rebuild counter
  int64_t* step_inc;          // device step counter advanced by k_keys (may be NULL)
  int32_t* gate_set;          // set to 1 by k_keys: the next stage-2 lists start valid (may be NULL)
  const double* pos[3];       // current positions
  const uint8_t* dead;        // per-row removal mark (may be NULL)
  const int32_t* slot;        // per-row logical slot (current)
  int32_t* key;               // per-row bucket
  int32_t* lin;               // per-row linear local cell
  const int32_t* gbucket;     // rank_off + linear local cell -> global bucket (cell_off + Morton rank)
  int32_t* count;             // per-bucket count
  int32_t* cursor;            // per-bucket scatter cursor
  int32_t* start;             // per-bucket first row (exclusive scan), nb+1 entries
  int32_t* tmp;               // bucket-ordered source rows
  int32_t* cell_row;          // destination row -> linear local cell
  CellRange* range;           // rank_off + linear local cell -> rows, written by k_scatter's range blocks
  int64_t nr;                 // entries of gbucket / range (all simulations)
  int32_t dims, periodic, full, geometric;
  unsigned long long* emax_clear;   // cleared by k_keys every step (may be NULL)
  int32_t* fix_clear;               // two counters cleared by k_keys every step (may be NULL)
  unsigned long long* efin_clear;   // two words cleared by k_keys when it rebuilds (may be NULL)
};

struct ScanArgs {
  const int32_t* flag;
  int32_t* count;
  int32_t* cursor;
  int32_t* start;
  int64_t n;
  unsigned long long* status;
  unsigned int* ticket;
  unsigned int* done;
  unsigned int* epoch;        // device launch epoch (>= 1)
  int32_t nblocks;
};

struct PermuteArgs {
  RebuildArgs rb;
  PermCols cols;
};

struct VerletArgs {
  RebuildArgs rb;
  PermCols cols;
  const double* apos[3];      // alternate (sorted) positions
  const int32_t* aslot;       // alternate (sorted) slots
  double* xref[3];
  int32_t* list;
  int32_t* list_n;
  int64_t ldl;
  int32_t cap;
  int32_t* overflow;
  double rl2;                 // largest s with sqrt(s) < list radius (exact test on d^2)
  int32_t build;              // 0: copy-back only (no list)
  int32_t nopermute;          // 1: rows stay in place; candidates via tmp, cells via lin
  int32_t lds_r;              // list entries a lane keeps sorted in LDS before spilling
  int32_t fixup;              // 1: only the rows of bricks k_verlet_brick left (brick_skip)
  BrickSrc bs;                // brick tiling (k_verlet_brick and its fixup)
};

// LDS budget of the list build: lds_r * block * 8 bytes.
constexpr int VERLET_LDS_BYTES = 24576;

struct XferArgs {
  RowGeom g;
  const int32_t* slot;
  void* field;          // device column
  void* staging;        // logical-order staging (per-sim n values, sim offset applied)
  int32_t bytes;
  int32_t to_device;
  int64_t n;            // values in staging
  int32_t sim;          // simulation whose rows are transferred
  int32_t pad;
};

// One field of every simulation gathered into slot order at
// out[row_off + slot] (evolve!(saveEach) snapshots); n_out[sim] = N.
struct SnapArgs {
  RowGeom g;
  const int32_t* slot;
  const uint64_t* field;
  uint64_t* out;        // NULL: the counts only
  int64_t* n_out;       // may be NULL
};

// One implicit ADI sweep along `axis` (the reference's tridiagonal solvers,
// src/integrators.jl:141-428): every line of interior nodes solved by the
// Thomas algorithm with constant diagonals; cst/mst hold c*[i] and
// m[i] = 1 / (main - sub c*[i-1]) along the axis, the same for every line.
struct AdiArgs {
  const double* src;
  double* dst;
  const double* cst;
  const double* mst;
  int32_t n[3];
  int32_t dims, axis;
  double sub, main, sup;
  int64_t lines;
};

struct ExportArgs {
  StageArgs s;
  const int32_t* slot;
  int32_t* counts;      // per logical slot
  int64_t* lists;       // per logical slot * cap
  int64_t* counts64;
  int64_t cap;
  double radius;
  int32_t take_all;     // Verlet: export the list itself
  int32_t* overflow;
  int32_t only_sim;
  int32_t clv;          // CellLinked on a skin list: exact walk in the cell order (clv_each)
  double e_bound;       // clv: bound on |x - x_ref| of every agent
};

// Adjoint of the overdamped-Euler mechanics step (no reference counterpart).
// Rows are not permuted while it runs, so lambda keeps its row mapping.
struct AdjointArgs {
  RowGeom g;
  const double* X[3];           // positions x_t (the state the step read)
  const double* lam[3];         // lambda_{t+1}
  double* lam_out[3];           // lambda_t
  const int32_t* list;
  const int32_t* list_n;
  int64_t ldl;
  const int32_t* slot;
  const double* target;         // sim s, save m, axis a, slot i: [tgt_base[s] + (m*D + a)*N_s + i]
  const int64_t* tgt_base;      // per-sim start of its targets
  int32_t save;                 // save index whose direct term enters lambda_t, or -1
  int32_t periodic;
  double dt;
  double scale;                 // 2 / (n_saves * N) per sim is applied with inv_n below
  const double* inv_n;          // per sim 1 / (n_saves * N)
  double* gpart;                // per block 3 running parameter-gradient sums
};

struct LossArgs {
  RowGeom g;
  const double* X[3];
  const int32_t* slot;
  const double* target;         // as in AdjointArgs
  const int64_t* tgt_base;
  int32_t save;
  int32_t dims;
  double* lpart;                // per block running loss sums
};

hipError_t launch_adjoint_step(int dims, int nblocks, int block, hipStream_t s, const AdjointArgs& a);
hipError_t launch_adjoint_seed(int dims, int nblocks, int block, hipStream_t s, const AdjointArgs& a);
hipError_t launch_loss(int nblocks, int block, hipStream_t s, const LossArgs& a);
hipError_t launch_disp_check(int dims, int nblocks, int block, hipStream_t s, const StageArgs& a);

hipError_t launch_medium_inject(hipStream_t s, double* c, unsigned long long* add, int64_t nodes);
hipError_t launch_medium_count(int nblocks, int block, hipStream_t s, const MediumArgs& a);
hipError_t launch_medium_step(hipStream_t s, const MediumArgs& a);
hipError_t launch_adi_sweep(hipStream_t s, const AdiArgs& a);
// boundary nodes of dst (any coordinate 0 or n-1) set from init
hipError_t launch_adi_border(hipStream_t s, double* dst, const double* init, const int32_t* n, int32_t dims,
                             int64_t nodes);

hipError_t launch_rule_decide(int nblocks, int block, hipStream_t s, const RuleArgs& a);
// copies positions, vol and the user parameters of the committed agents into
// the rule's neighbor view (RuleArgs::nb_*; dst = those pointers)
hipError_t launch_rule_nb_copy(int nblocks, int block, hipStream_t s, const RuleArgs& a, double* const* dst);
hipError_t launch_rule_apply(int nblocks, int block, hipStream_t s, const RuleArgs& a);
hipError_t launch_compact_holes(int nblocks, int block, hipStream_t s, const RuleArgs& a);
hipError_t launch_compact_slots(int nblocks, int block, hipStream_t s, const RuleArgs& a);
hipError_t launch_commit_counts(hipStream_t s, const RuleArgs& a, int32_t nsims);

// ------------------------------------------------------------- launchers
// 16 Ki counts per tile: the look-back chain (one tile after another) sets
// the scan's time, so fewer, larger tiles are faster
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;
inline int scan_blocks(int64_t n) { return (int)((n + SCAN_TILE - 1) / SCAN_TILE); }

// NB_CELL launches one block per brick (nbricks) with brick_lds bytes of LDS
hipError_t launch_stage(int dims, int nb, int law, int nblocks, int block, hipStream_t s, const StageArgs& a,
                        int64_t nbricks, size_t brick_lds);
hipError_t launch_keys(int nblocks, int block, hipStream_t s, const RebuildArgs& a);
hipError_t launch_scan(hipStream_t s, const ScanArgs& a);
// the row scatter, and with ranges the cell range table (one launch)
hipError_t launch_scatter(int nblocks, int block, hipStream_t s, const RebuildArgs& a, bool ranges);
hipError_t launch_permute(int nblocks, int block, hipStream_t s, const PermuteArgs& a);
hipError_t launch_verlet(int dims, int nblocks, int block, hipStream_t s, const VerletArgs& a);
// Verlet list build on LDS brick tiles (one block per brick), to be followed
// by launch_verlet with fixup = 1 for the bricks too dense for LDS.
hipError_t launch_verlet_brick(int dims, int periodic, int64_t nbricks, int block, size_t lds, hipStream_t s,
                               const VerletArgs& a);
hipError_t launch_set_flag(hipStream_t s, int32_t* flag, int32_t v);
hipError_t launch_tick(hipStream_t s, int64_t* counter);
hipError_t launch_xfer(int nblocks, int block, hipStream_t s, const XferArgs& a);
hipError_t launch_snapshot(int nblocks, int block, hipStream_t s, const SnapArgs& a);
hipError_t launch_export(int dims, int nb, int nblocks, int block, hipStream_t s, const ExportArgs& e);
hipError_t launch_clcell(int dims, int nblocks, int block, hipStream_t s, const StageArgs& a);

}  // namespace cbm

#endif
