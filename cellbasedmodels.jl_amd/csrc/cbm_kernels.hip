// cbm_kernels.hip -- gfx950 kernels of the per-step agent update.
//
//   k_stage<D,NB,LAW>  neighbor gather + pair law + one integrator stage
//                      (the generated agentODE kernel of
//                      src/AgentStructure/functionDE.jl:92-108 fused with the
//                      CBMIntegrators broadcasts, src/integrators.jl:465-734)
//   k_brick<D,LAW>     the same stage for the cell-linked walk on LDS tiles
//                      of Morton-aligned cell bricks (cellPosNeigh order,
//                      src/neighbors.jl:641-700)
//   k_keys             cell key + bucket histogram   (assignCells, src/neighbors.jl:763-821)
//   k_scan             single-pass exclusive scan    (cumsum, src/neighbors.jl:934-950)
//   k_scatter          bucket scatter                (sortAgentsInCells, src/neighbors.jl:837-899)
//                      and, in the same launch, per-cell row ranges from the bucket starts
//   k_permute          deterministic in-cell order + column permutation
//   k_verlet           Verlet list build from cells (src/neighbors.jl:92-165, 1071-1183)
//
// Agents are stored sorted by (cell, logical slot); every neighbor gather
// therefore reads a compact, mostly L2-resident window of rows, and the
// candidate order (cells in cellPosNeigh order, ascending slot within a
// cell, Verlet lists ascending by slot) equals the reference CPU loop order.
#include "cbm_internal.h"
#include "cbm_device.h"
#include "cbm_stage.h"
#include "cbm_medium.h"

namespace cbm {

// Axis cell of a position: cellPos (src/neighbors.jl:636-638) on the padded
// grid, or the periodic grid of n cells of edge L/n.
// geometric: bin by position then clamp (Verlet list builds, which must keep
// every pair within the list radius in adjacent cells even for agents just
// outside the box, where the reference clamping would separate them).
__device__ __forceinline__ int32_t axis_cell(const SimDev& sd, int a, double x, int periodic, int geometric) {
  if (geometric) {
    if (x < sd.lo[a]) return 0;
    int64_t c = floor_div(x - sd.lo[a], sd.gedge[a]) + (periodic ? 0 : 1);
    return (int32_t)(c > sd.gn[a] - 1 ? sd.gn[a] - 1 : c);
  }
  if (periodic) {
    int64_t c = floor_div(x - sd.lo[a], sd.gedge[a]);
    if (c < 0) c = 0;
    if (c > sd.gn[a] - 1) c = sd.gn[a] - 1;
    return (int32_t)c;
  }
  if (x > sd.hi[a]) return sd.gn[a] - 1;
  if (x < sd.lo[a]) return 0;
  return (int32_t)(floor_div(x - sd.lo[a], sd.gedge[a]) + 1);
}

template <int D, int NB, int LAW, int PER, int NOISE>
__global__ __launch_bounds__(256) void k_stage(StageArgs a) {
  stage_body<D, NB, PairLaw<LAW>, PER, NOISE>(a);
}

// a stage of the 2-D rod model (no periodic box, no noise)
template <int NB>
__global__ __launch_bounds__(256) void k_rods(StageArgs a) {
  rod_stage_body<NB>(a);
}

// cell-linked stage on LDS brick tiles (dynamic LDS: brick_lds_bytes)
template <int D, int LAW, int PER, int NOISE>
__global__ __launch_bounds__(256) void k_brick(StageArgs a) {
  extern __shared__ unsigned char brick_lds[];
  brick_body<D, PairLaw<LAW>, PER, NOISE>(a, brick_lds);
}

// list stage with the pair work flattened over the block (dynamic LDS:
// flat_lds_bytes)
template <int D, int LAW, int PER, int NOISE>
__global__ __launch_bounds__(256) void k_flat(StageArgs a) {
  extern __shared__ unsigned char flat_lds[];
  flat_body<D, PairLaw<LAW>, PER, NOISE>(a, flat_lds);
}

// CellLinked on a skin list: the step's ordered lists (dynamic LDS: the
// per-lane key queues, blockDim x lcap uint16), the stages over them, and
// the exact walk of the agents they deferred
template <int D, int LAW, int PER, int NOISE>
__global__ __launch_bounds__(256) void k_clv(StageArgs a) {
  extern __shared__ unsigned char clv_lds[];
  clv_stage_body<D, PairLaw<LAW>, PER, NOISE>(a, clv_lds);
}

template <int D, int LAW, int PER, int NOISE>
__global__ __launch_bounds__(256) void k_clv_fix(StageArgs a) {
  if (a.tick && blockIdx.x == 0 && threadIdx.x == 0) *a.tick += 1;   // the step's last kernel
  if (a.fix_n[a.fix_q] == 0) return;
  clv_fix_body<D, PairLaw<LAW>, PER, NOISE>(a);
}

// --------------------------------------------------------- rebuild chain

__global__ __launch_bounds__(256) void k_keys(RebuildArgs a) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // runs whether or not this step rebuilds: advance the device step and
    // clear the flag the next step's final stage will set
    if (a.step_inc) *a.step_inc += 1;
    if (a.flag_clear) *a.flag_clear = 0;
    if (a.emax_clear) *a.emax_clear = 0ull;
    if (a.fix_clear) { a.fix_clear[0] = 0; a.fix_clear[1] = 0; }
    if (a.gate_set) *a.gate_set = 1;
  }
  if (a.flag && *a.flag == 0) return;
  const RowIdx ri = row_index(a.g);
  const SimDev& sd = a.g.sims[ri.sim];
  const SimState& st = a.g.st[ri.sim];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (a.rebuilds) *a.rebuilds += 1;
    if (a.efin_clear) { a.efin_clear[0] = 0ull; a.efin_clear[1] = 0ull; }
  }
  if (ri.local == 0) a.count[sd.cell_off + sd.ncells + 1] = (int32_t)(sd.n_max - st.n_tot);
  if (ri.local >= st.n_tot) return;
  const int64_t r = ri.row;
  int32_t c, lin = 0;
  if (a.dead && a.dead[r]) c = (int32_t)(sd.cell_off + sd.ncells);
  else if (a.full) c = (int32_t)sd.cell_off;
  else {
    int32_t q[3] = {0, 0, 0};
    for (int d = 0; d < a.dims; ++d) q[d] = axis_cell(sd, d, a.pos[d][r], a.periodic, a.geometric);
    lin = q[0] + sd.gn[0] * (q[1] + sd.gn[1] * q[2]);
    c = a.gbucket[sd.rank_off + lin];        // cell_off + Morton rank of the cell
  }
  a.key[r] = c;
  a.lin[r] = lin;
  atomicAdd(&a.count[c], 1);
}

// Single-pass exclusive scan with decoupled look-back (Merrill & Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back", 2016).  Tiles
// are ordered by an atomic ticket, so a tile only ever waits on tiles that
// are already running; status words carry a launch epoch kept in device
// memory (so graph replays stay distinct) and never need clearing.  The scan also zeroes count[] and cursor[] for the next
// rebuild.

__device__ __forceinline__ unsigned long long scan_word(unsigned int epoch, unsigned int flag, unsigned int value) {
  return ((unsigned long long)(epoch & 0x3FFFFFFFu) << 34) | ((unsigned long long)flag << 32) | value;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan(ScanArgs a) {
  if (a.flag && *a.flag == 0) return;
  __shared__ unsigned int s_tile;
  __shared__ int32_t s_wave[SCAN_THREADS / 64];
  __shared__ int32_t s_prefix;
  const int tid = threadIdx.x;
  __shared__ unsigned int s_epoch;
  if (tid == 0) {
    s_tile = atomicAdd(a.ticket, 1u);
    s_epoch = __hip_atomic_load(a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t base = tile * SCAN_TILE + (int64_t)tid * SCAN_ITEMS;
  // a thread's items are contiguous: whole chunks move as 16-byte vectors
  const bool whole = base + SCAN_ITEMS <= a.n;
  int32_t v[SCAN_ITEMS];
  int32_t sum = 0;
  if (whole) {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k += 4) {
      const int4 q = *(const int4*)(a.count + base + k);
      v[k] = q.x; v[k + 1] = q.y; v[k + 2] = q.z; v[k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) v[k] = base + k < a.n ? a.count[base + k] : 0;
  }
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) sum += v[k];
  // wavefront inclusive scan of per-thread sums
  const int lane = tid & 63, wave = tid >> 6;
  int32_t incl = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) s_wave[wave] = incl;
  __syncthreads();
  int32_t wave_off = 0, total = 0;
#pragma unroll
  for (int w = 0; w < SCAN_THREADS / 64; ++w) {
    const int32_t t = s_wave[w];
    if (w < wave) wave_off += t;
    total += t;
  }
  if (wave == 0) {
    // Publish this tile's aggregate, then look back 64 tiles at a time: lane
    // l reads tile p - l; the window ends at the nearest tile that already
    // holds its inclusive prefix (tiles before 0 count as an inclusive 0).
    const unsigned int ep = s_epoch & 0x3FFFFFFFu;
    int32_t prefix = 0;
    if (lane == 0)
      __hip_atomic_store(&a.status[tile], scan_word(ep, tile == 0 ? 2u : 1u, (unsigned int)total),
                         __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    int64_t p = tile - 1;
    while (p >= 0) {
      const int64_t q = p - lane;
      unsigned int fl = 2u, val = 0u;
      if (q >= 0) {
        const unsigned long long w = __hip_atomic_load(&a.status[q], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        fl = ((unsigned int)(w >> 34) == ep) ? ((unsigned int)(w >> 32) & 3u) : 0u;
        val = (unsigned int)w;
      }
      const unsigned long long inc = __ballot(fl == 2u);
      const int stop = inc ? __ffsll((unsigned long long)inc) - 1 : 63;
      // every tile up to the stop lane must have published; else wait
      const unsigned long long unready = __ballot(fl == 0u && lane <= stop);
      if (unready) { __builtin_amdgcn_s_sleep(1); continue; }
      int32_t x = lane <= stop ? (int32_t)val : 0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
      prefix += x;
      if (inc) break;
      p -= 64;
    }
    if (lane == 0) {
      if (tile > 0)
        __hip_atomic_store(&a.status[tile], scan_word(ep, 2u, (unsigned int)(prefix + total)),
                           __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      s_prefix = prefix;
    }
  }
  __syncthreads();
  int32_t run = s_prefix + wave_off + (incl - sum);
  if (whole) {
    const int4 z = make_int4(0, 0, 0, 0);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k += 4) {
      int4 q;
      q.x = run; run += v[k];
      q.y = run; run += v[k + 1];
      q.z = run; run += v[k + 2];
      q.w = run; run += v[k + 3];
      *(int4*)(a.start + base + k) = q;
      *(int4*)(a.count + base + k) = z;
      if (a.cursor) *(int4*)(a.cursor + base + k) = z;
    }
    if (base + SCAN_ITEMS == a.n) a.start[a.n] = run;
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      const int64_t idx = base + k;
      if (idx < a.n) {
        a.start[idx] = run;
        a.count[idx] = 0;
        if (a.cursor) a.cursor[idx] = 0;
        run += v[k];
        if (idx == a.n - 1) a.start[a.n] = run;
      }
    }
  }
  if (tid == 0) {
    const unsigned int fin = atomicAdd(a.done, 1u);
    if (fin == (unsigned int)a.nblocks - 1u) {
      // every tile has read the epoch and published; advance it for the next launch
      const unsigned int nx = (s_epoch + 1u) & 0x3FFFFFFFu;
      __hip_atomic_store(a.epoch, nx == 0u ? 1u : nx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(a.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(a.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// One launch after the scan: blocks [0, nrow) scatter the rows into bucket
// order; blocks past them gather the rows [s, e) of every grid cell by linear
// cell (all simulations) from the bucket starts, written in table order so
// the stores coalesce.
__global__ __launch_bounds__(256) void k_scatter(RebuildArgs a, int nrow) {
  if (a.flag && *a.flag == 0) return;
  if ((int)blockIdx.x >= nrow) {
    const int64_t g = (int64_t)(blockIdx.x - nrow) * blockDim.x + threadIdx.x;
    if (g >= a.nr) return;
    const int32_t b = a.gbucket[g];
    a.range[g] = CellRange{a.start[b], a.start[b + 1]};
    return;
  }
  const RowIdx ri = row_index_of(a.g, (int)blockIdx.x, nrow);
  const SimDev& sd = a.g.sims[ri.sim];
  if (ri.local >= a.g.st[ri.sim].n_tot) return;
  const int64_t r = ri.row;
  const int32_t c = a.key[r];
  int32_t pos;
  if (a.full && c == (int32_t)sd.cell_off) pos = a.start[c] + a.slot[r];
  else pos = a.start[c] + atomicAdd(&a.cursor[c], 1);
  a.tmp[pos] = (int32_t)r;
}


// Each source row finds its rank by logical slot inside its bucket and
// moves every column to that row of the alternate buffers.
__global__ __launch_bounds__(256) void k_permute(PermuteArgs pa) {
  const RebuildArgs& a = pa.rb;
  if (a.flag && *a.flag == 0) return;
  const RowIdx ri = row_index(a.g);
  const SimDev& sd = a.g.sims[ri.sim];
  if (ri.local >= a.g.st[ri.sim].n_tot) return;
  const int64_t r = ri.row;
  const int32_t c = a.key[r];
  if (c == (int32_t)(sd.cell_off + sd.ncells)) return;   // removed agent
  const int32_t s0 = a.start[c];
  int32_t dst;
  const int32_t my = a.slot[r];
  if (a.full) dst = s0 + my;
  else {
    const int32_t s1 = a.start[c + 1];
    int32_t rank = 0;
    for (int32_t q = s0; q < s1; ++q) rank += (a.slot[a.tmp[q]] < my) ? 1 : 0;
    dst = s0 + rank;
  }
  for (int k = 0; k < pa.cols.n; ++k) {
    if (pa.cols.bytes[k] == 8)
      ((uint64_t*)pa.cols.alt[k])[dst] = ((const uint64_t*)pa.cols.cur[k])[r];
    else
      ((uint32_t*)pa.cols.alt[k])[dst] = ((const uint32_t*)pa.cols.cur[k])[r];
  }
  a.cell_row[dst] = a.lin[r];
}

// Verlet list build on the freshly sorted rows (alternate buffers), then
// copy-back of every column to the current buffers and snapshot of the
// reference positions.  Entries are global rows ordered by ascending slot,
// the order of the reference brute-force build (src/neighbors.jl:114-126).

// Accept row j into row i's list.  The first R entries are kept sorted by
// slot in this lane's LDS buffer (entry q at buf[q * B]); a list longer than
// R is spilled to the global list once and continued there by insertion.
__device__ __forceinline__ void verlet_accept(const VerletArgs& va, int64_t i, int2* __restrict__ buf,
                                              int B, int R, int32_t& cnt, int32_t j, int32_t sj) {
  if (cnt < R) {
    int32_t pos = cnt;
    while (pos > 0) {
      const int2 prev = buf[(pos - 1) * B];
      if (prev.x < sj) break;
      buf[pos * B] = prev;
      --pos;
    }
    buf[pos * B] = make_int2(sj, j);
  } else {
    if (cnt == R)
      for (int q = 0; q < R; ++q) va.list[(int64_t)q * va.ldl + i] = buf[q * B].y;
    if (cnt < va.cap) {
      int32_t pos = cnt;
      while (pos > 0) {
        const int32_t prev = va.list[(int64_t)(pos - 1) * va.ldl + i];
        if (va.aslot[prev] < sj) break;
        va.list[(int64_t)pos * va.ldl + i] = prev;
        --pos;
      }
      va.list[(int64_t)pos * va.ldl + i] = j;
    } else {
      *va.overflow = 1;
    }
  }
  ++cnt;
}

template <int D>
__device__ __forceinline__ int32_t verlet_row(const VerletArgs& va, const SimDev& sd, int64_t i, int2* buf, int B) {
  const RebuildArgs& a = va.rb;
  if (!va.build) {
    for (int k = 0; k < va.cols.n; ++k) {
      if (va.cols.bytes[k] == 8)
        ((uint64_t*)va.cols.cur[k])[i] = ((const uint64_t*)va.cols.alt[k])[i];
      else
        ((uint32_t*)va.cols.cur[k])[i] = ((const uint32_t*)va.cols.alt[k])[i];
    }
    return 0;
  }
  double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < D; ++c) xi[c] = va.apos[c][i];
  const int32_t c = va.nopermute ? a.lin[i] : a.cell_row[i];
  const int32_t cc[3] = {c % sd.gn[0], (c / sd.gn[0]) % sd.gn[1], c / (sd.gn[0] * sd.gn[1])};
  const CellRange* __restrict__ cr = a.range + sd.rank_off;
  const int R = va.lds_r;
  int32_t cnt = 0;
  // Columns of three x-adjacent cells; the candidates of a column are walked
  // as one flattened range so that lanes of a wavefront stay in step.  The
  // cell ranges of the next column are loaded before the current column's
  // candidates are walked.
  constexpr int NCOL = D == 1 ? 1 : (D == 2 ? 3 : 9);
  auto col_ranges = [&](int col, CellRange* r3) {
    int32_t yz = 0;
    bool ok = true;
    if (D > 1) {
      int32_t q[2] = {cc[1] + col % 3 - 1, cc[2] + col / 3 - 1};
#pragma unroll
      for (int e = 0; e < D - 1; ++e) {
        int32_t v = q[e];
        if (a.periodic) v = v < 0 ? v + sd.gn[1 + e] : (v >= sd.gn[1 + e] ? v - sd.gn[1 + e] : v);
        else ok = ok && v >= 0 && v < sd.gn[1 + e];
        q[e] = v;
      }
      yz = sd.gn[0] * (q[0] + (D > 2 ? sd.gn[1] * q[1] : 0));
    }
#pragma unroll
    for (int dxc = 0; dxc < 3; ++dxc) {
      int32_t v = cc[0] + dxc - 1;
      bool okx = ok;
      if (a.periodic) v = v < 0 ? v + sd.gn[0] : (v >= sd.gn[0] ? v - sd.gn[0] : v);
      else okx = okx && v >= 0 && v < sd.gn[0];
      r3[dxc] = CellRange{0, 0};
      if (okx) r3[dxc] = cr[yz + v];
    }
  };
  CellRange nxt[3];
  col_ranges(0, nxt);
#pragma unroll
  for (int col = 0; col < NCOL; ++col) {
    int32_t s3[3], n3[3];
#pragma unroll
    for (int dxc = 0; dxc < 3; ++dxc) { s3[dxc] = nxt[dxc].s; n3[dxc] = nxt[dxc].e - nxt[dxc].s; }
    if (col + 1 < NCOL) col_ranges(col + 1, nxt);
    // candidates in chunks of VC whose loads are issued together; slots past
    // the end load nothing and are skipped
    constexpr int VC = 2;
    const int32_t n01 = n3[0] + n3[1], tot = n01 + n3[2];
    for (int32_t t0 = 0; t0 < tot; t0 += VC) {
      int32_t jj[VC];
#pragma unroll
      for (int u = 0; u < VC; ++u) {
        const int32_t t = t0 + u;
        jj[u] = (int32_t)i;
        if (t < tot) {
          const int32_t jq = t < n3[0] ? s3[0] + t : (t < n01 ? s3[1] + (t - n3[0]) : s3[2] + (t - n01));
          jj[u] = va.nopermute ? a.tmp[jq] : jq;
        }
      }
      double px[VC][3];
      int32_t sl[VC];
#pragma unroll
      for (int u = 0; u < VC; ++u) {
#pragma unroll
        for (int q = 0; q < D; ++q) px[u][q] = xi[q];
        sl[u] = 0;
        if (t0 + u < tot) {
#pragma unroll
          for (int q = 0; q < D; ++q) px[u][q] = va.apos[q][jj[u]];
          sl[u] = va.aslot[jj[u]];
        }
      }
#pragma unroll
      for (int u = 0; u < VC; ++u) {
        if (t0 + u >= tot || jj[u] == i) continue;
        double dx[3];
#pragma unroll
        for (int q = 0; q < D; ++q) {
          dx[q] = xi[q] - px[u][q];
          if (a.periodic) dx[q] = min_image(dx[q], sd.L[q]);
        }
        double s2 = dx[0] * dx[0];
        if (D > 1) s2 = s2 + dx[1] * dx[1];
        if (D > 2) s2 = s2 + dx[2] * dx[2];
        // s2 <= rl2  <=>  sqrt(s2) < rl for a correctly rounded sqrt
        if (!(s2 <= va.rl2)) continue;
        verlet_accept(va, i, buf, B, R, cnt, jj[u], sl[u]);
      }
    }
  }
  if (cnt <= R)
    for (int32_t q = 0; q < cnt; ++q) va.list[(int64_t)q * va.ldl + i] = buf[q * B].y;
  const int32_t kept = cnt < va.cap ? cnt : va.cap;
  va.list_n[i] = kept;
  for (int k = 0; k < (va.nopermute ? 0 : va.cols.n); ++k) {
    if (va.cols.bytes[k] == 8)
      ((uint64_t*)va.cols.cur[k])[i] = ((const uint64_t*)va.cols.alt[k])[i];
    else
      ((uint32_t*)va.cols.cur[k])[i] = ((const uint32_t*)va.cols.alt[k])[i];
  }
#pragma unroll
  for (int q = 0; q < D; ++q) va.xref[q][i] = xi[q];
  return kept;
}


template <int D>
__global__ __launch_bounds__(256) void k_verlet(VerletArgs va) {
  extern __shared__ int2 vbuf[];
  const RebuildArgs& a = va.rb;
  if (a.flag && *a.flag == 0) return;
  if (va.fixup) {
    // only blocks holding rows of bricks too dense for k_verlet_brick
    const int b = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    if (!va.bs.blk_flag[b]) return;
    __syncthreads();
    if (threadIdx.x == 0) va.bs.blk_flag[b] = 0;
  }
  const RowIdx ri = row_index(a.g);
  const SimDev& sd = a.g.sims[ri.sim];
  const int64_t N = a.g.st[ri.sim].N;
  if (va.fixup) {
    if (ri.local >= sd.n_max || !va.bs.brick_skip[ri.row]) return;
    va.bs.brick_skip[ri.row] = 0;
  }
  if (ri.local >= N) return;
  (void)verlet_row<D>(va, sd, ri.row, vbuf + threadIdx.x, (int)blockDim.x);
}

// The list build on LDS brick tiles: the halo of a brick (positions and
// slots of the sorted rows) is staged once; each home agent walks its 3^D
// cells there, queues the candidates within the list radius (the same exact
// d^2 test as k_verlet) in a per-lane LDS queue, sorts them by slot in
// registers and writes them as rows.  A list longer than the queue is rebuilt by insertion into
// the global list, as k_verlet does past its LDS buffer, so the lists equal
// k_verlet's.  Also copies the sorted rows back and snapshots x_ref.
template <int D, int PER>
__global__ __launch_bounds__(256) void k_verlet_brick(VerletArgs va) {
  extern __shared__ unsigned char vlds[];
  const RebuildArgs& a = va.rb;
  if (a.flag && *a.flag == 0) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int cap = va.bs.cap;
  const int H = brick_halo_cells(D, va.bs.bdim[0], D > 1 ? va.bs.bdim[1] : 1, D > 2 ? va.bs.bdim[2] : 1);
  BrickTile t;
  t.hdr = (int32_t*)vlds;
  double* pos[3] = {nullptr, nullptr, nullptr};
  unsigned char* p = vlds + 4 * BRICK_HDR;
#pragma unroll
  for (int c = 0; c < D; ++c) { pos[c] = (double*)p; p += (size_t)8 * cap; }
  int32_t* slot = (int32_t*)p;
  p += (size_t)4 * cap;
  t.hs = (int32_t*)p;
  p += (size_t)4 * (H + 1);
  t.hrow = (int32_t*)p;
  p += (size_t)4 * H;
  uint16_t* lst = (uint16_t*)p;       // lane queue, entry u at [u * nt + tid], sorted by slot
  p += (size_t)2 * VBRICK_Q * BRICK_MAX_BLOCK;
  t.owner = (uint16_t*)p;
  p += (size_t)2 * cap;
  t.home_li = (uint16_t*)p;

  BrickCols cols;
#pragma unroll
  for (int c = 0; c < 3; ++c) { cols.S[c] = va.apos[c]; cols.pos[c] = pos[c]; }
  cols.xd = nullptr;
  cols.xd_lds = nullptr;
  cols.xi = va.aslot;
  cols.xi_lds = slot;
  double unused = 0.0;
  const int b = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  if (brick_stage<D, PER, 4, false, BX_I32>(va.bs, b, t, cols, unused) != 1) return;

  const SimDev& sd = a.g.sims[t.sim];
  const int64_t N = a.g.st[t.sim].N;
  double L[3], hL[3];
  box_halves<D, PER>(sd, L, hL);
  auto row_of = [&](int32_t q) {
    const int h = t.owner[q] & 0x7FFF;
    return (int64_t)t.hrow[h] + (q - t.hs[h]);
  };
  const int32_t nh = t.r1 - t.r0;
  for (int32_t k = tid; k < nh; k += nt) {
    const int32_t li = t.home_li[k];
    const int h = t.owner[li] & 0x7FFF;
    const int64_t i = row_of(li);
    if (i - sd.row_off >= N) continue;
    double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < D; ++c) xi[c] = pos[c][li];
    auto within = [&](int32_t q) {
      double dx[3];
#pragma unroll
      for (int c = 0; c < D; ++c) {
        dx[c] = xi[c] - pos[c][q];
        if (PER) dx[c] = wrap_dx(dx[c], L[c], hL[c]);
      }
      double s2 = dx[0] * dx[0];
      if (D > 1) s2 = s2 + dx[1] * dx[1];
      if (D > 2) s2 = s2 + dx[2] * dx[2];
      return s2 <= va.rl2 && q != li;   // sqrt(s2) < r_list, not the agent itself
    };
    int32_t nk = 0;
    brick_walk<D>(t, h, [&](int32_t q) {
      if (!within(q)) return;
      if (nk < VBRICK_Q) lst[nk * nt + tid] = (uint16_t)q;
      ++nk;
    });
    int32_t kept;
    if (nk <= VBRICK_Q) {
      // (slot, LDS index) keys; unused entries sort last
      int64_t key[VBRICK_Q];
#pragma unroll
      for (int u = 0; u < VBRICK_Q; ++u) {
        const int32_t q = u < nk ? lst[u * nt + tid] : 0;
        key[u] = u < nk ? ((int64_t)slot[q] << 16) | q : INT64_MAX;
      }
      sort_net<VBRICK_Q>(key);
      kept = nk < va.cap ? nk : va.cap;
#pragma unroll
      for (int u = 0; u < VBRICK_Q; ++u)
        if (u < kept) va.list[(int64_t)u * va.ldl + i] = (int32_t)row_of((int32_t)(key[u] & 0xFFFF));
      if (nk > va.cap) *va.overflow = 1;
    } else {
      int32_t cnt = 0;
      brick_walk<D>(t, h, [&](int32_t q) {
        if (within(q)) verlet_accept(va, i, nullptr, 0, 0, cnt, (int32_t)row_of(q), slot[q]);
      });
      kept = cnt < va.cap ? cnt : va.cap;
    }
    va.list_n[i] = kept;
    for (int c = 0; c < va.cols.n; ++c) {
      if (va.cols.bytes[c] == 8)
        ((uint64_t*)va.cols.cur[c])[i] = ((const uint64_t*)va.cols.alt[c])[i];
      else
        ((uint32_t*)va.cols.cur[c])[i] = ((const uint32_t*)va.cols.alt[c])[i];
    }
#pragma unroll
    for (int c = 0; c < D; ++c) va.xref[c][i] = xi[c];
  }
}


// ------------------------------------------------------------------ rules
// Built-in agentRule bodies (src/AgentStructure/functionRule.jl:9-111 for
// @addAgent/@removeAgent).  Decisions are pure functions of the agent's
// committed state and Philox(seed, step, id, sim, purpose), so the decide and
// apply passes agree and the result is independent of thread order.
struct RuleDecision { int32_t add; int32_t dead; };

// A daughter's user parameters: the parent's (@addAgent copies every
// parameter it is not given, src/AgentStructure/functionRule.jl:46-55), in
// both buffers, since user DE variables flip with the positions.
__device__ __forceinline__ void rule_copy_u(const RuleArgs& a, int64_t r, int64_t dst) {
  if (!a.u[0]) return;
#pragma unroll
  for (int q = 0; q < USER_PARAMS; ++q) { const double v = a.u[q][r]; a.u[q][dst] = v; a.u_alt[q][dst] = v; }
}

__device__ __forceinline__ RuleDecision rule_decide(const RuleArgs& a, int sim, int64_t r, int64_t step) {
  RuleDecision d{0, 0};
  const int64_t id = a.id[r];
  if (a.rule == 2) {                       // parity KAT: odd ids removed, even ids add one
    if (id % 2 == 1) d.dead = 1; else d.add = 1;
  } else if (a.rule == 3) {                // death with probability rate*dt
    const u4 q = draw(a.seed, step, id, sim, 0x400u);
    d.dead = u53(q.x, q.y) < a.rp[0] * a.dt ? 1 : 0;
  } else if (a.rule == 1) {                // growth and division
    d.add = (a.vol[r] >= a.rp[3] * a.vb[r]) ? 1 : 0;
  } else if (a.rule == 5) {                // rods2dGrowth: two daughters replace a rod past lTarget
    const bool div = a.rcol[RC_LEN][r] > a.rcol[RC_LT][r];
    d.add = div ? 2 : 0;
    d.dead = div ? 1 : 0;
  }
  return d;
}

__global__ __launch_bounds__(256) void k_rule_decide(RuleArgs a) {
  const RowIdx ri = row_index(a.g);
  const int64_t N = a.g.st[ri.sim].N;
  if (ri.local >= N) return;
  const int64_t r = ri.row;
  const RuleDecision d = rule_decide(a, ri.sim, r, *a.step_ptr);
  const int32_t s = a.slot[r];
  a.n_add_slot[s] = d.add;
  a.dead_slot[s] = d.dead;
}

// Re-evaluates the rule and writes its effects: parent updates in place,
// daughters at rows N + (exclusive prefix of daughters by slot), which is the
// reference's single-thread append order; new ids idMax + 1 + that prefix.
__global__ __launch_bounds__(256) void k_rule_apply(RuleArgs a) {
  const RowIdx ri = row_index(a.g);
  SimState& st = a.g.st[ri.sim];
  const SimDev& sd = a.g.sims[ri.sim];
  const int64_t N = st.N;
  const int64_t nmax = sd.n_max;
  const int64_t A = a.off_add[nmax];
  if (ri.local == 0) {
    st.n_add = A;
    st.n_remove = a.off_dead[nmax];
    st.n_tot = (N + A <= nmax) ? N + A : nmax;
    if (N + A > nmax) *a.err = 1;          // reference D12 drops the agent silently
  }
  if (ri.local >= N) return;
  const int64_t r = ri.row;
  const int64_t step = *a.step_ptr;
  const RuleDecision d = rule_decide(a, ri.sim, r, step);
  const int32_t s = a.slot[r];
  a.dead[r] = (uint8_t)d.dead;
  const int D = a.dims;
  if (a.rule == 1) {                        // write-new values for every agent
    a.gr_new[r] = a.gr[r];
    a.vb_new[r] = a.vb[r];
  }
  if (a.rule == 2) {                        // x = x + 1 (test/testCommunity.jl:247)
    const double xold = a.pos[0][r];
    a.pos[0][r] = xold + 1.0;
    if (d.add) {
      const int64_t o = a.off_add[s];
      const int64_t nw = N + o;
      if (nw >= nmax) return;
      const int64_t dst = sd.row_off + nw;
      for (int c = 0; c < D; ++c) {
        const double v = (c == 0) ? xold : a.pos[c][r];
        a.pos[c][dst] = v;
        a.pos_alt[c][dst] = v;
      }
      if (a.lagged) for (int c = 0; c < D; ++c) { a.F[c][dst] = a.F[c][r]; a.Fnew[c][dst] = a.F[c][r]; }
      a.id[dst] = st.id_max + 1 + o;
      a.lf[dst] = a.id[r];
      rule_copy_u(a, r, dst);
      a.slot[dst] = (int32_t)nw;
      a.dead[dst] = 0;
    }
  } else if (a.rule == 5 && d.add) {
    // src/models/rods.jl:141-160: ww = uniform(-0.1, 0.1); each daughter at
    // +-(l + d)/4 along the axis, length (l - d)/2, omega +- ww and its own
    // lTarget = uniform(lMax - sigma, lMax + sigma); the rest copied
    const int64_t o = a.off_add[s];
    const int64_t id = a.id[r];
    const u4 q0 = draw(a.seed, step, id, ri.sim, 0x700u);
    const u4 q1 = draw(a.seed, step, id, ri.sim, 0x701u);
    const double ww = (0.1 - (-0.1)) * u53(q0.x, q0.y) + (-0.1);
    const double lo = a.rp[2] - a.rp[1], hi = a.rp[2] + a.rp[1];
    const double lt[2] = {(hi - lo) * u53(q0.z, q0.w) + lo, (hi - lo) * u53(q1.x, q1.y) + lo};
    const double x = a.pos[0][r], y = a.pos[1][r];
    const double th = a.rcol[RC_TH][r], om = a.rcol[RC_OM][r];
    const double l = a.rcol[RC_LEN][r], dd = a.rcol[RC_DIAM][r];
    const double q = (l + dd) / 4;
    const double cth = cos(th), sth = sin(th);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int64_t nw = N + o + k;
      if (nw >= nmax) continue;
      const int64_t dst = sd.row_off + nw;
#pragma unroll
      for (int j = 0; j < RC_N; ++j) { const double v = a.rcol[j][r]; a.rcol[j][dst] = v; a.rcol_alt[j][dst] = v; }
      const double qs = k ? -q : q;
      const double xn = qs * cth + x, yn = qs * sth + y;
      a.pos[0][dst] = xn; a.pos_alt[0][dst] = xn;
      a.pos[1][dst] = yn; a.pos_alt[1][dst] = yn;
      const double ln = (l - dd) / 2, on = k ? om - ww : om + ww;
      a.rcol[RC_LEN][dst] = ln; a.rcol_alt[RC_LEN][dst] = ln;
      a.rcol[RC_OM][dst] = on; a.rcol_alt[RC_OM][dst] = on;
      a.rcol[RC_LT][dst] = lt[k]; a.rcol_alt[RC_LT][dst] = lt[k];
      a.id[dst] = st.id_max + 1 + o + k;
      a.lf[dst] = a.lf[r];
      a.slot[dst] = (int32_t)nw;
      a.dead[dst] = 0;
    }
  } else if (a.rule == 1 && d.add) {
    const int64_t o = a.off_add[s];
    const int64_t nw = N + o;
    if (nw >= nmax) return;
    const int64_t dst = sd.row_off + nw;
    const int64_t id = a.id[r];
    const u4 ra = draw(a.seed, step, id, ri.sim, 0x200u);
    const u4 rg = draw(a.seed, step, id, ri.sim, 0x300u);
    double ax[3] = {0.0, 0.0, 0.0};
    if (D == 1) ax[0] = (u53(ra.x, ra.y) < 0.5) ? -1.0 : 1.0;
    else if (D == 2) {
      const double th = 6.283185307179586 * u53(ra.x, ra.y);
      ax[0] = cos(th); ax[1] = sin(th);
    } else {
      const double z = 2.0 * u53(ra.x, ra.y) - 1.0;
      const double ph = 6.283185307179586 * u53(ra.z, ra.w);
      const double rr = sqrt(1.0 - z * z);
      ax[0] = rr * cos(ph); ax[1] = rr * sin(ph); ax[2] = z;
    }
    const double split = a.rp[2];
    const double half = a.vol[r] / 2.0;
    const double gm = a.rp[0], gs = a.rp[1];
    const double g1 = gm * (1.0 + gs * (2.0 * u53(rg.x, rg.y) - 1.0));
    const double g2 = gm * (1.0 + gs * (2.0 * u53(rg.z, rg.w) - 1.0));
    for (int c = 0; c < D; ++c) {
      const double x = a.pos[c][r];
      double xn = x + split * ax[c];
      if (a.periodic) xn = wrap_box(xn, sd.lo[c], sd.L[c]);
      a.pos[c][dst] = xn;
      a.pos_alt[c][dst] = xn;
      a.pos[c][r] = x - split * ax[c];
    }
    if (a.lagged) for (int c = 0; c < D; ++c) { a.F[c][dst] = a.F[c][r]; a.Fnew[c][dst] = a.F[c][r]; }
    a.vol[r] = half; a.vol[dst] = half; a.vol_alt[dst] = half;
    a.vb_new[r] = half; a.vb_new[dst] = half; ((double*)a.vb)[dst] = half;
    a.gr_new[r] = g1; a.gr_new[dst] = g2; ((double*)a.gr)[dst] = g2;
    a.id[dst] = st.id_max + 1 + o;
    a.lf[dst] = a.lf[r];
    rule_copy_u(a, r, dst);
    a.slot[dst] = (int32_t)nw;
    a.dead[dst] = 0;
  }
}

// update! compaction (src/CommunityStructure/update.jl:15-119) in the order
// of the reference's serial run: holes ascending, survivors taken from the
// tail; hole m receives survivor m whenever hole < survivor.  With
// N' = N + A - R this is: alive slots below N' stay, the alive slots at or
// above N' move, the m-th from the top into the m-th hole below N'.
__global__ __launch_bounds__(256) void k_compact_holes(RuleArgs a) {
  const RowIdx ri = row_index(a.g);
  const SimState& st = a.g.st[ri.sim];
  if (ri.local >= st.n_tot) return;
  const int64_t r = ri.row;
  if (!a.dead[r]) return;
  const int64_t Np = st.N + st.n_add - st.n_remove;
  const int32_t s = a.slot[r];
  if (s < Np) a.holes[a.off_dead[s]] = s;
}

__global__ __launch_bounds__(256) void k_compact_slots(RuleArgs a) {
  const RowIdx ri = row_index(a.g);
  const SimState& st = a.g.st[ri.sim];
  const int64_t N = st.N, A = st.n_add, R = st.n_remove;
  const int64_t Np = N + A - R;
  if (ri.local < st.n_tot) {
    const int64_t r = ri.row;
    if (!a.dead[r]) {
      const int64_t s = a.slot[r];
      if (s >= Np) {
        const int64_t m = (N + A - 1 - s) - (R - (int64_t)a.off_dead[s]);
        a.slot[r] = a.holes[m];
      }
    }
  }
}

// ----------------------------------------------------------------- adjoint
// Block-wide sum of NV doubles into out[tile*NV + q], tile = the block's
// row-geometry index (running sums: each tile always owns the same entries,
// so the totals are deterministic).
template <int NV>
__device__ __forceinline__ void block_accumulate(double* v, double* out) {
  const int64_t tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  __shared__ double red[NV][256 / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double x = v[q];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) red[q][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += red[q][w];
      out[tile * NV + q] += s;
    }
  }
}

// lambda_t = lambda_{t+1} + dt sum_j J_ij (lambda_i - lambda_j) + dL/dx_t,
// J = phi I + (phi'/d) r r^T for r = x_i - x_j: the transpose of the Euler
// step's Jacobian gathered over the (symmetric) Verlet list.  Parameter
// gradient: dt lambda_{t+1,i} . sum_j dphi/dtheta r.
template <int D>
__global__ __launch_bounds__(256) void k_adjoint_step(AdjointArgs a) {
  const RowIdx ri = row_index(a.g);
  const SimDev& sd = a.g.sims[ri.sim];
  const int64_t N = a.g.st[ri.sim].N;
  const bool live = ri.local < N;
  double gth[3] = {0.0, 0.0, 0.0};
  if (live) {
    const int64_t i = ri.row;
    const RepAdhGrad law(sd.p);
    double xi[3] = {0.0, 0.0, 0.0}, li[3] = {0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < D; ++c) { xi[c] = a.X[c][i]; li[c] = a.lam[c][i]; }
    double L[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < D; ++c) L[c] = sd.L[c];
    auto pair = [&](const double* xj, const double* lj, bool in) {
      double r[3], u[3];
#pragma unroll
      for (int c = 0; c < D; ++c) {
        r[c] = xi[c] - xj[c];
        if (a.periodic) r[c] = r[c] > 0.5 * L[c] ? r[c] - L[c] : (r[c] < -0.5 * L[c] ? r[c] + L[c] : r[c]);
        u[c] = li[c] - lj[c];
      }
      double s2 = r[0] * r[0];
      if (D > 1) s2 += r[1] * r[1];
      if (D > 2) s2 += r[2] * r[2];
      const double d = sqrt(s2);
      const double id = 1.0 / d;
      double phi, dphid, dth[3];
      const bool on = law(d, id, phi, dphid, dth) && in;
      double ru = 0.0, lr = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) { ru += r[c] * u[c]; lr += li[c] * r[c]; }
      const double pe = on ? phi : 0.0;
      const double ge = on ? dphid * ru : 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) v[c] += pe * u[c] + ge * r[c];
#pragma unroll
      for (int q2 = 0; q2 < 3; ++q2) gth[q2] += (on ? dth[q2] : 0.0) * lr;
    };
    // chunks of two list entries, next indices prefetched; entries past the
    // count take the agent itself without a load (d = 0, never active)
    const int32_t cnt = a.list_n[i];
    const int64_t ldl = a.ldl;
    const int32_t* __restrict__ lp = a.list + i;
    int32_t jn[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) jn[u] = (u < cnt) ? lp[(int64_t)u * ldl] : (int32_t)i;
    for (int32_t q0 = 0; q0 < cnt; q0 += 2) {
      double xj[2][3], lj[2][3];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
#pragma unroll
        for (int c = 0; c < D; ++c) { xj[u][c] = xi[c]; lj[u][c] = li[c]; }
        if (q0 + u < cnt) {
#pragma unroll
          for (int c = 0; c < D; ++c) { xj[u][c] = a.X[c][jn[u]]; lj[u][c] = a.lam[c][jn[u]]; }
        }
      }
      lp += 2 * ldl;
#pragma unroll
      for (int u = 0; u < 2; ++u) jn[u] = (q0 + 2 + u < cnt) ? lp[(int64_t)u * ldl] : (int32_t)i;
#pragma unroll
      for (int u = 0; u < 2; ++u) pair(xj[u], lj[u], q0 + u < cnt);
    }
    double direct[3] = {0.0, 0.0, 0.0};
    if (a.save >= 0) {
      const double* tg = a.target + a.tgt_base[ri.sim] + a.slot[i];
      const double w = 2.0 * a.inv_n[ri.sim];
#pragma unroll
      for (int c = 0; c < D; ++c)
        direct[c] = w * (xi[c] - tg[((int64_t)a.save * D + c) * N]);
    }
#pragma unroll
    for (int c = 0; c < D; ++c) a.lam_out[c][i] = li[c] + a.dt * v[c] + direct[c];
#pragma unroll
    for (int q2 = 0; q2 < 3; ++q2) gth[q2] *= a.dt;
  }
  block_accumulate<3>(gth, a.gpart);
}

// lambda_T = dL/dx_T (the last save), or 0.
template <int D>
__global__ __launch_bounds__(256) void k_adjoint_seed(AdjointArgs a) {
  const RowIdx ri = row_index(a.g);
  if (ri.local >= a.g.st[ri.sim].N) return;
  const int64_t i = ri.row;
  for (int c = 0; c < D; ++c) {
    double v = 0.0;
    if (a.save >= 0) {
      const int64_t N = a.g.st[ri.sim].N;
      const double* tg = a.target + a.tgt_base[ri.sim] + a.slot[i];
      v = 2.0 * a.inv_n[ri.sim] * (a.X[c][i] - tg[((int64_t)a.save * D + c) * N]);
    }
    a.lam_out[c][i] = v;
  }
}

// Running per-block sums of |x - x*|^2 at a save point.
__global__ __launch_bounds__(256) void k_loss(LossArgs a) {
  const RowIdx ri = row_index(a.g);
  double e = 0.0;
  if (ri.local < a.g.st[ri.sim].N) {
    const int64_t i = ri.row;
    const int64_t N = a.g.st[ri.sim].N;
    const double* tg = a.target + a.tgt_base[ri.sim] + a.slot[i];
    for (int c = 0; c < a.dims; ++c) {
      const double d = a.X[c][i] - tg[((int64_t)a.save * a.dims + c) * N];
      e += d * d;
    }
  }
  block_accumulate<1>(&e, a.lpart);
}

// Verlet validity check of positions X (in S) against the list snapshot.
template <int D>
__global__ __launch_bounds__(256) void k_disp_check(StageArgs a) {
  const RowIdx ri = row_index(a.g);
  if (ri.local >= a.g.st[ri.sim].N) return;
  const SimDev& sd = a.g.sims[ri.sim];
  const int64_t i = ri.row;
  double s2 = 0.0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    double dx = a.S[c][i] - a.xref[c][i];
    if (a.periodic) dx = min_image(dx, sd.L[c]);
    s2 = (c == 0) ? dx * dx : s2 + dx * dx;
  }
  if (s2 > a.thr2) *a.flag = 1;   // sqrt(s2) >= thr
}

// ------------------------------------------------------------------ medium
// Agents per medium node: cellInMesh (src/metrics.jl:18-20) per axis, the
// nearest-lower node of a cell-centred grid, clamped to the grid.  Integer
// atomics make the count independent of thread order; the reference's
// per-agent read-modify-write of c is racy.
__global__ __launch_bounds__(256) void k_medium_count(MediumArgs m) {
  const RowIdx ri = row_index(m.g);
  if (ri.sim != 0 || ri.local >= m.g.st[0].N) return;
  const int64_t r = ri.row;
  int64_t q[3] = {0, 0, 0};
  for (int a = 0; a < m.dims; ++a) {
    const double x = m.pos[a][r];
    int64_t v;
    if (x > m.hi[a]) v = m.n[a] - 1;
    else if (x < m.lo[a]) v = 0;
    else v = floor_div(x - m.lo[a], m.dx[a]);
    q[a] = v > m.n[a] - 1 ? m.n[a] - 1 : v;
  }
  atomicAdd(&m.count[q[0] + (int64_t)m.n[0] * (q[1] + (int64_t)m.n[1] * q[2])], 1);
}

// The rules' medium additions of the step (32.32 fixed point per node) land
// in the field before its update; the sums are cleared for the next step.
__global__ __launch_bounds__(256) void k_medium_inject(double* c, unsigned long long* add, int64_t nodes) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nodes) return;
  const long long v = (long long)add[k];
  if (v != 0) {
    c[k] = c[k] + (double)v * (1.0 / MEDIUM_ADD_SCALE);
    add[k] = 0;
  }
}

// FTCS: c' = c + dt (D sum_a (c[+a] - 2c + c[-a])/dx_a^2 - (uptake c) n_agents),
// ghost nodes mirroring the boundary node (zero flux, conserves mass).
__global__ __launch_bounds__(256) void k_medium_step(MediumArgs m) { medium_step_body<BuiltinMedium>(m); }

// DGADI sweep: one thread per line of interior nodes (lines over the other
// axes' interior indices), forward elimination into dst, then back
// substitution in place, in the reference's order of operations.
__global__ __launch_bounds__(256) void k_adi_sweep(AdiArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.lines) return;
  const int64_t stride[3] = {1, a.n[0], (int64_t)a.n[0] * a.n[1]};
  int64_t base = 0, rem = t;
  for (int b = 0; b < a.dims; ++b) {
    if (b == a.axis) continue;
    const int64_t m = a.n[b] - 2;
    base += (1 + rem % m) * stride[b];
    rem /= m;
  }
  const int64_t s = stride[a.axis];
  const int n = a.n[a.axis];
  const double* x = a.src + base;
  double* y = a.dst + base;
  double d = (x[s] - a.sub * x[0]) / a.main;
  y[s] = d;
  for (int i = 2; i <= n - 3; ++i) {
    d = (x[i * s] - a.sub * d) * a.mst[i];
    y[i * s] = d;
  }
  d = ((x[(int64_t)(n - 2) * s] - a.sup * x[(int64_t)(n - 1) * s]) - a.sub * d) * a.mst[n - 2];
  y[(int64_t)(n - 2) * s] = d;
  for (int i = n - 3; i >= 1; --i) {
    d = y[i * s] - a.cst[i] * d;
    y[i * s] = d;
  }
}

struct BorderArgs { double* dst; const double* init; int32_t n[3]; int32_t dims; int64_t nodes; };

__global__ __launch_bounds__(256) void k_adi_border(BorderArgs a) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.nodes) return;
  const int64_t nx = a.n[0], ny = a.n[1];
  const int64_t p[3] = {k % nx, (k / nx) % ny, k / (nx * ny)};
  bool edge = false;
  for (int b = 0; b < a.dims; ++b) edge = edge || p[b] == 0 || p[b] == a.n[b] - 1;
  if (edge) a.dst[k] = a.init[k];
}

// N' = N + NAdd_ - NRemove_ and idMax_ += NAdd_ (update.jl:204-214), after
// every compaction kernel has read the old counts; flags the neighbor
// rebuild when the population changed.
__global__ void k_commit_counts(RuleArgs a, int32_t nsims) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsims) return;
  SimState& st = a.g.st[s];
  const int64_t A = st.n_add, R = st.n_remove;
  st.N = st.N + A - R;
  st.id_max += A;
  if (A + R > 0) *a.rebuild_flag = 1;
}

// --------------------------------------------------------------- utilities
__global__ void k_set_flag(int32_t* flag, int32_t v) { if (threadIdx.x == 0 && blockIdx.x == 0) *flag = v; }
__global__ void k_tick(int64_t* counter) { if (threadIdx.x == 0 && blockIdx.x == 0) *counter += 1; }


// Upload scatters logical-order values into rows; download gathers them back.
__global__ __launch_bounds__(256) void k_xfer(XferArgs a) {
  const RowIdx ri = row_index(a.g);
  if (ri.sim != a.sim) return;
  const int64_t r = ri.row;
  const int64_t s = a.slot[r];
  if (ri.local >= a.g.st[ri.sim].N) return;   // committed agents only
  if (s >= a.n) return;
  if (a.bytes == 8) {
    if (a.to_device) ((uint64_t*)a.field)[r] = ((const uint64_t*)a.staging)[s];
    else ((uint64_t*)a.staging)[s] = ((const uint64_t*)a.field)[r];
  } else {
    if (a.to_device) ((uint32_t*)a.field)[r] = ((const uint32_t*)a.staging)[s];
    else ((uint32_t*)a.staging)[s] = ((const uint32_t*)a.field)[r];
  }
}

// saveRAM! (src/CommunityStructure/IO.jl:6-31) on the device: the committed
// agents' values of one field in slot order, every simulation at once.
__global__ __launch_bounds__(256) void k_snapshot(SnapArgs a) {
  const RowIdx ri = row_index(a.g);
  const SimState& st = a.g.st[ri.sim];
  if (a.n_out && ri.local == 0) a.n_out[ri.sim] = st.N;
  if (!a.out || ri.local >= st.N) return;
  a.out[a.g.sims[ri.sim].row_off + a.slot[ri.row]] = a.field[ri.row];
}

// Neighbor-set export (mode 0: lists in slots, mode 1: counts only) for the
// sorted-set parity checks and the neighbor-count agentRule of the
// reference KAT (test/testCommunity.jl:361-415).

template <int D, int NB>
__global__ __launch_bounds__(256) void k_export(ExportArgs e) {
  const StageArgs& a = e.s;
  const RowIdx ri = row_index(a.g);
  if (ri.sim != e.only_sim) return;
  const SimDev& sd = a.g.sims[ri.sim];
  const int64_t N = a.g.st[ri.sim].N;
  if (ri.local >= N) return;
  const int64_t i = ri.row;
  const int64_t si = e.slot[i];
  double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < D; ++c) xi[c] = a.S[c][i];
  int64_t k = 0;
  for_candidates<D, NB>(a, sd, N, i, [&](int64_t j) {
    bool take = e.take_all != 0;
    if (!take) {
      double dx[3];
#pragma unroll
      for (int c = 0; c < D; ++c) {
        dx[c] = xi[c] - a.S[c][j];
        if (a.periodic) dx[c] = min_image(dx[c], sd.L[c]);
      }
      double s2 = dx[0] * dx[0];
      if (D > 1) s2 = s2 + dx[1] * dx[1];
      if (D > 2) s2 = s2 + dx[2] * dx[2];
      take = sqrt(s2) < e.radius;
    }
    if (take) {
      if (e.lists) {
        if (k < e.cap) e.lists[si * e.cap + k] = e.slot[j];
        else *e.overflow = 1;
      }
      ++k;
    }
  });
  if (e.counts) e.counts[si] = (int32_t)k;
  if (e.counts64) e.counts64[si] = k;
}

// The export of a CellLinked model run on a skin list: every candidate of the
// model's cells within the radius, in its order (clv_each).
template <int D, int PER>
__global__ __launch_bounds__(256) void k_export_clv(ExportArgs e) {
  const StageArgs& a = e.s;
  const RowIdx ri = row_index(a.g);
  if (ri.sim != e.only_sim) return;
  const SimDev& sd = a.g.sims[ri.sim];
  const int64_t N = a.g.st[ri.sim].N;
  if (ri.local >= N) return;
  const int64_t i = ri.row;
  const int64_t si = e.slot[i];
  double xi[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < D; ++c) xi[c] = a.S[c][i];
  int64_t k = 0;
  clv_each<D, PER>(a, sd, i, xi, reach2_bound(e.radius), e.radius + e.e_bound, [&](int64_t j, double s2) {
    if (!(sqrt(s2) < e.radius)) return;
    if (e.lists) {
      if (k < e.cap) e.lists[si * e.cap + k] = e.slot[j];
      else *e.overflow = 1;
    }
    ++k;
  });
  if (e.counts) e.counts[si] = (int32_t)k;
  if (e.counts64) e.counts64[si] = k;
}

// Packed cells (cl_pack) of every row's position a.S into a.ccell_out.
template <int D>
__global__ __launch_bounds__(256) void k_clcell(StageArgs a) {
  const RowIdx ri = row_index(a.g);
  const SimDev& sd = a.g.sims[ri.sim];
  if (ri.local >= a.g.st[ri.sim].n_tot) return;
  double x[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < D; ++c) x[c] = a.S[c][ri.row];
  a.ccell_out[ri.row] = cl_pack<D>(sd, x, a.periodic);
}

// ------------------------------------------------------------- launchers
struct StageLaunch { dim3 grid, bgrid, block; size_t lds; hipStream_t s; };   // bgrid: one block per brick

template <int D, int NB, int PER, int NOISE>
static hipError_t launch_stage_law(int law, const StageLaunch& L, const StageArgs& a) {
  if (NB == NB_CELL) {
    // bricks that fit in LDS, then the rows of those that did not
    switch (law) {
      case LAW_LINEAR: hipLaunchKernelGGL((k_brick<D, LAW_LINEAR, PER, NOISE>), L.bgrid, L.block, L.lds, L.s, a); break;
      case LAW_REP_ADH: hipLaunchKernelGGL((k_brick<D, LAW_REP_ADH, PER, NOISE>), L.bgrid, L.block, L.lds, L.s, a); break;
      case LAW_AGGREGATION: hipLaunchKernelGGL((k_brick<D, LAW_AGGREGATION, PER, NOISE>), L.bgrid, L.block, L.lds, L.s, a); break;
      case LAW_SOFT_SPHERES: hipLaunchKernelGGL((k_brick<D, LAW_SOFT_SPHERES, PER, NOISE>), L.bgrid, L.block, L.lds, L.s, a); break;
      default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (a.brick2 && a.nbricks2 > 0) {
      // the sub-bricks of the bricks that did not fit
      StageArgs b = a;
      b.brick = a.brick2;
      for (int q = 0; q < 3; ++q) b.bdim[q] = a.bdim2[q];
      b.brick_cap = a.brick_cap2;
      b.brick_sub = 1;
      const dim3 g2((unsigned)a.nbricks2);
      const size_t lds2 = (size_t)a.brick_lds2;
      switch (law) {
        case LAW_LINEAR: hipLaunchKernelGGL((k_brick<D, LAW_LINEAR, PER, NOISE>), g2, L.block, lds2, L.s, b); break;
        case LAW_REP_ADH: hipLaunchKernelGGL((k_brick<D, LAW_REP_ADH, PER, NOISE>), g2, L.block, lds2, L.s, b); break;
        case LAW_AGGREGATION: hipLaunchKernelGGL((k_brick<D, LAW_AGGREGATION, PER, NOISE>), g2, L.block, lds2, L.s, b); break;
        default: hipLaunchKernelGGL((k_brick<D, LAW_SOFT_SPHERES, PER, NOISE>), g2, L.block, lds2, L.s, b); break;
      }
      e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  }
  switch (law) {
    case LAW_LINEAR: hipLaunchKernelGGL((k_stage<D, NB, LAW_LINEAR, PER, NOISE>), L.grid, L.block, 0, L.s, a); break;
    case LAW_REP_ADH: hipLaunchKernelGGL((k_stage<D, NB, LAW_REP_ADH, PER, NOISE>), L.grid, L.block, 0, L.s, a); break;
    case LAW_AGGREGATION: hipLaunchKernelGGL((k_stage<D, NB, LAW_AGGREGATION, PER, NOISE>), L.grid, L.block, 0, L.s, a); break;
    case LAW_SOFT_SPHERES: hipLaunchKernelGGL((k_stage<D, NB, LAW_SOFT_SPHERES, PER, NOISE>), L.grid, L.block, 0, L.s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// a stage of CellLinked on a skin list, then the exact walk of the agents it
// deferred (a grid of at most 256 blocks, one per CU; it exits at once when none were)
template <int D, int PER, int NOISE>
static hipError_t launch_clv_stage(int law, const StageLaunch& L, const StageArgs& a) {
  const dim3 fg(L.grid.x < 256 ? L.grid.x : 256);
  // the ordering stage keeps lists over 16 entries in LDS (16-bit keys)
  const size_t lds = a.clv_order ? (size_t)L.block.x * (size_t)a.lcap * 2 : 0;
  switch (law) {
    case LAW_LINEAR:
      hipLaunchKernelGGL((k_clv<D, LAW_LINEAR, PER, NOISE>), L.grid, L.block, lds, L.s, a);
      hipLaunchKernelGGL((k_clv_fix<D, LAW_LINEAR, PER, NOISE>), fg, L.block, 0, L.s, a);
      break;
    case LAW_REP_ADH:
      hipLaunchKernelGGL((k_clv<D, LAW_REP_ADH, PER, NOISE>), L.grid, L.block, lds, L.s, a);
      hipLaunchKernelGGL((k_clv_fix<D, LAW_REP_ADH, PER, NOISE>), fg, L.block, 0, L.s, a);
      break;
    case LAW_AGGREGATION:
      hipLaunchKernelGGL((k_clv<D, LAW_AGGREGATION, PER, NOISE>), L.grid, L.block, lds, L.s, a);
      hipLaunchKernelGGL((k_clv_fix<D, LAW_AGGREGATION, PER, NOISE>), fg, L.block, 0, L.s, a);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int D, int PER, int NOISE>
static hipError_t launch_flat(int law, const StageLaunch& L, const StageArgs& a) {
  const size_t lds = (size_t)flat_lds_bytes(D, a.fchunk, a.fcap);
  switch (law) {
    case LAW_LINEAR: hipLaunchKernelGGL((k_flat<D, LAW_LINEAR, PER, NOISE>), L.grid, L.block, lds, L.s, a); break;
    case LAW_REP_ADH: hipLaunchKernelGGL((k_flat<D, LAW_REP_ADH, PER, NOISE>), L.grid, L.block, lds, L.s, a); break;
    case LAW_AGGREGATION: hipLaunchKernelGGL((k_flat<D, LAW_AGGREGATION, PER, NOISE>), L.grid, L.block, lds, L.s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int D, int PER, int NOISE>
static hipError_t launch_stage_nb(int nb, int law, const StageLaunch& L, const StageArgs& a) {
  const bool reach_law = law == LAW_LINEAR || law == LAW_REP_ADH || law == LAW_AGGREGATION;
  if (nb == NB_LIST && a.clv && reach_law) return launch_clv_stage<D, PER, NOISE>(law, L, a);
  if (nb == NB_LIST && a.fchunk > 0 && reach_law) return launch_flat<D, PER, NOISE>(law, L, a);
  switch (nb) {
    case NB_FULL: return launch_stage_law<D, NB_FULL, PER, NOISE>(law, L, a);
    case NB_CELL: return launch_stage_law<D, NB_CELL, PER, NOISE>(law, L, a);
    default: return launch_stage_law<D, NB_LIST, PER, NOISE>(law, L, a);
  }
}

template <int D, int NOISE>
static hipError_t launch_stage_dims(int nb, int law, int periodic, const StageLaunch& L, const StageArgs& a) {
  if (law == LAW_RODS) {      // 2-D, no periodic box, no noise: the host refuses the rest
    if (D != 2 || NOISE || periodic) return hipErrorInvalidValue;
    switch (nb) {
      case NB_FULL: hipLaunchKernelGGL((k_rods<NB_FULL>), L.grid, L.block, 0, L.s, a); break;
      case NB_CELL: hipLaunchKernelGGL((k_rods<NB_CELL>), L.grid, L.block, 0, L.s, a); break;
      default: hipLaunchKernelGGL((k_rods<NB_LIST>), L.grid, L.block, 0, L.s, a); break;
    }
    return hipGetLastError();
  }
  if (law == LAW_NONE) {      // no interaction body: no neighbor walk at all
    hipLaunchKernelGGL((k_stage<D, NB_FULL, LAW_NONE, 0, NOISE>), L.grid, L.block, 0, L.s, a);
    return hipGetLastError();
  }
  return periodic ? launch_stage_nb<D, 1, NOISE>(nb, law, L, a) : launch_stage_nb<D, 0, NOISE>(nb, law, L, a);
}

hipError_t launch_stage(int dims, int nb, int law, int nblocks, int block, hipStream_t s, const StageArgs& a,
                        int64_t nbricks, size_t brick_lds) {
  StageLaunch L;
  L.grid = dim3((unsigned)nblocks);
  L.bgrid = dim3((unsigned)nbricks);
  L.block = dim3(block);
  L.lds = brick_lds;
  L.s = s;
  const bool noise = a.kind == ST_EM || a.kind == ST_EH1 || a.kind == ST_EH2;
  switch (dims) {
    case 1: return noise ? launch_stage_dims<1, 1>(nb, law, a.periodic, L, a) : launch_stage_dims<1, 0>(nb, law, a.periodic, L, a);
    case 2: return noise ? launch_stage_dims<2, 1>(nb, law, a.periodic, L, a) : launch_stage_dims<2, 0>(nb, law, a.periodic, L, a);
    default: return noise ? launch_stage_dims<3, 1>(nb, law, a.periodic, L, a) : launch_stage_dims<3, 0>(nb, law, a.periodic, L, a);
  }
}

hipError_t launch_keys(int nblocks, int block, hipStream_t s, const RebuildArgs& a) {
  hipLaunchKernelGGL(k_keys, dim3(nblocks), dim3(block), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_scan(hipStream_t s, const ScanArgs& a) {
  hipLaunchKernelGGL(k_scan, dim3(a.nblocks), dim3(SCAN_THREADS), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_scatter(int nblocks, int block, hipStream_t s, const RebuildArgs& a, bool ranges) {
  const int64_t nrb = ranges ? (a.nr + block - 1) / block : 0;
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)(nblocks + nrb)), dim3(block), 0, s, a, nblocks);
  return hipGetLastError();
}

hipError_t launch_permute(int nblocks, int block, hipStream_t s, const PermuteArgs& a) {
  hipLaunchKernelGGL(k_permute, dim3(nblocks), dim3(block), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_verlet(int dims, int nblocks, int block, hipStream_t s, const VerletArgs& a) {
  const size_t lds = a.build ? (size_t)a.lds_r * (size_t)block * sizeof(int2) : 0;
  switch (dims) {
    case 1: hipLaunchKernelGGL(k_verlet<1>, dim3(nblocks), dim3(block), lds, s, a); break;
    case 2: hipLaunchKernelGGL(k_verlet<2>, dim3(nblocks), dim3(block), lds, s, a); break;
    default: hipLaunchKernelGGL(k_verlet<3>, dim3(nblocks), dim3(block), lds, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_verlet_brick(int dims, int periodic, int64_t nbricks, int block, size_t lds, hipStream_t s,
                               const VerletArgs& a) {
  const dim3 g((unsigned)nbricks), bl(block);
  switch (dims * 2 + (periodic ? 1 : 0)) {
    case 2: hipLaunchKernelGGL((k_verlet_brick<1, 0>), g, bl, lds, s, a); break;
    case 3: hipLaunchKernelGGL((k_verlet_brick<1, 1>), g, bl, lds, s, a); break;
    case 4: hipLaunchKernelGGL((k_verlet_brick<2, 0>), g, bl, lds, s, a); break;
    case 5: hipLaunchKernelGGL((k_verlet_brick<2, 1>), g, bl, lds, s, a); break;
    case 6: hipLaunchKernelGGL((k_verlet_brick<3, 0>), g, bl, lds, s, a); break;
    default: hipLaunchKernelGGL((k_verlet_brick<3, 1>), g, bl, lds, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_adjoint_step(int dims, int nblocks, int block, hipStream_t s, const AdjointArgs& a) {
  switch (dims) {
    case 1: hipLaunchKernelGGL(k_adjoint_step<1>, dim3(nblocks), dim3(block), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_adjoint_step<2>, dim3(nblocks), dim3(block), 0, s, a); break;
    default: hipLaunchKernelGGL(k_adjoint_step<3>, dim3(nblocks), dim3(block), 0, s, a); break;
  }
  return hipGetLastError();
}
hipError_t launch_adjoint_seed(int dims, int nblocks, int block, hipStream_t s, const AdjointArgs& a) {
  switch (dims) {
    case 1: hipLaunchKernelGGL(k_adjoint_seed<1>, dim3(nblocks), dim3(block), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_adjoint_seed<2>, dim3(nblocks), dim3(block), 0, s, a); break;
    default: hipLaunchKernelGGL(k_adjoint_seed<3>, dim3(nblocks), dim3(block), 0, s, a); break;
  }
  return hipGetLastError();
}
hipError_t launch_loss(int nblocks, int block, hipStream_t s, const LossArgs& a) {
  hipLaunchKernelGGL(k_loss, dim3(nblocks), dim3(block), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_disp_check(int dims, int nblocks, int block, hipStream_t s, const StageArgs& a) {
  switch (dims) {
    case 1: hipLaunchKernelGGL(k_disp_check<1>, dim3(nblocks), dim3(block), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_disp_check<2>, dim3(nblocks), dim3(block), 0, s, a); break;
    default: hipLaunchKernelGGL(k_disp_check<3>, dim3(nblocks), dim3(block), 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_medium_count(int nblocks, int block, hipStream_t s, const MediumArgs& a) {
  hipLaunchKernelGGL(k_medium_count, dim3(nblocks), dim3(block), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_medium_inject(hipStream_t s, double* c, unsigned long long* add, int64_t nodes) {
  hipLaunchKernelGGL(k_medium_inject, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, s, c, add, nodes);
  return hipGetLastError();
}
hipError_t launch_medium_step(hipStream_t s, const MediumArgs& a) {
  hipLaunchKernelGGL(k_medium_step, dim3((unsigned)((a.nodes + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_adi_sweep(hipStream_t s, const AdiArgs& a) {
  hipLaunchKernelGGL(k_adi_sweep, dim3((unsigned)((a.lines + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_adi_border(hipStream_t s, double* dst, const double* init, const int32_t* n, int32_t dims,
                             int64_t nodes) {
  BorderArgs a;
  a.dst = dst; a.init = init; a.dims = dims; a.nodes = nodes;
  for (int b = 0; b < 3; ++b) a.n[b] = n[b];
  hipLaunchKernelGGL(k_adi_border, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

struct NbCopy { double* dst[4 + USER_PARAMS]; };

__global__ __launch_bounds__(256) void k_rule_nb_copy(RuleArgs a, NbCopy d) {
  const RowIdx ri = row_index(a.g);
  if (ri.local >= a.g.st[ri.sim].N) return;
  const int64_t r = ri.row;
  for (int c = 0; c < a.dims; ++c) d.dst[c][r] = a.pos[c][r];
  d.dst[3][r] = a.vol[r];
#pragma unroll
  for (int k = 0; k < USER_PARAMS; ++k) d.dst[4 + k][r] = a.u[k][r];
}

hipError_t launch_rule_nb_copy(int nblocks, int block, hipStream_t s, const RuleArgs& a, double* const* dst) {
  NbCopy d;
  for (int k = 0; k < 4 + USER_PARAMS; ++k) d.dst[k] = dst[k];
  hipLaunchKernelGGL(k_rule_nb_copy, dim3(nblocks), dim3(block), 0, s, a, d);
  return hipGetLastError();
}

hipError_t launch_rule_decide(int nblocks, int block, hipStream_t s, const RuleArgs& a) {
  hipLaunchKernelGGL(k_rule_decide, dim3(nblocks), dim3(block), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_rule_apply(int nblocks, int block, hipStream_t s, const RuleArgs& a) {
  hipLaunchKernelGGL(k_rule_apply, dim3(nblocks), dim3(block), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_compact_holes(int nblocks, int block, hipStream_t s, const RuleArgs& a) {
  hipLaunchKernelGGL(k_compact_holes, dim3(nblocks), dim3(block), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_compact_slots(int nblocks, int block, hipStream_t s, const RuleArgs& a) {
  hipLaunchKernelGGL(k_compact_slots, dim3(nblocks), dim3(block), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_commit_counts(hipStream_t s, const RuleArgs& a, int32_t nsims) {
  hipLaunchKernelGGL(k_commit_counts, dim3((nsims + 63) / 64), dim3(64), 0, s, a, nsims);
  return hipGetLastError();
}

hipError_t launch_set_flag(hipStream_t s, int32_t* flag, int32_t v) {
  hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, s, flag, v);
  return hipGetLastError();
}

hipError_t launch_tick(hipStream_t s, int64_t* counter) {
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, s, counter);
  return hipGetLastError();
}

hipError_t launch_xfer(int nblocks, int block, hipStream_t s, const XferArgs& a) {
  hipLaunchKernelGGL(k_xfer, dim3(nblocks), dim3(block), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_snapshot(int nblocks, int block, hipStream_t s, const SnapArgs& a) {
  hipLaunchKernelGGL(k_snapshot, dim3(nblocks), dim3(block), 0, s, a);
  return hipGetLastError();
}

template <int D>
static hipError_t launch_export_nb(int nb, dim3 grid, dim3 block, hipStream_t s, const ExportArgs& e) {
  switch (nb) {
    case NB_FULL: hipLaunchKernelGGL((k_export<D, NB_FULL>), grid, block, 0, s, e); break;
    case NB_CELL: hipLaunchKernelGGL((k_export<D, NB_CELL>), grid, block, 0, s, e); break;
    default: hipLaunchKernelGGL((k_export<D, NB_LIST>), grid, block, 0, s, e); break;
  }
  return hipGetLastError();
}

hipError_t launch_clcell(int dims, int nblocks, int block, hipStream_t s, const StageArgs& a) {
  switch (dims) {
    case 1: hipLaunchKernelGGL(k_clcell<1>, dim3(nblocks), dim3(block), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_clcell<2>, dim3(nblocks), dim3(block), 0, s, a); break;
    default: hipLaunchKernelGGL(k_clcell<3>, dim3(nblocks), dim3(block), 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_export(int dims, int nb, int nblocks, int block, hipStream_t s, const ExportArgs& e) {
  dim3 grid(nblocks), blk(block);
  if (e.clv) {
    switch (dims * 2 + (e.s.periodic ? 1 : 0)) {
      case 2: hipLaunchKernelGGL((k_export_clv<1, 0>), grid, blk, 0, s, e); break;
      case 3: hipLaunchKernelGGL((k_export_clv<1, 1>), grid, blk, 0, s, e); break;
      case 4: hipLaunchKernelGGL((k_export_clv<2, 0>), grid, blk, 0, s, e); break;
      case 5: hipLaunchKernelGGL((k_export_clv<2, 1>), grid, blk, 0, s, e); break;
      case 6: hipLaunchKernelGGL((k_export_clv<3, 0>), grid, blk, 0, s, e); break;
      default: hipLaunchKernelGGL((k_export_clv<3, 1>), grid, blk, 0, s, e); break;
    }
    return hipGetLastError();
  }
  switch (dims) {
    case 1: return launch_export_nb<1>(nb, grid, blk, s, e);
    case 2: return launch_export_nb<2>(nb, grid, blk, s, e);
    default: return launch_export_nb<3>(nb, grid, blk, s, e);
  }
}

}  // namespace cbm
