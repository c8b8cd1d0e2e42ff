// kmpm_kernels.hip -- gfx950 kernels of the explicit MPM step that do not
// depend on the tile shape (cell keys, counting sort, tile lists, dt,
// observables), and the launchers that select the tile variant.
//
// Step structure (USL):
//   sort   : cell key per particle -> stable counting sort by tile-major
//            cell key (ties keep the previous step's sorted order, which is
//            ascending tag after upload: a deterministic permutation)
//   P2G    : one workgroup per active tile; lanes = stencil nodes; each lane
//            accumulates the particles of a cell in registers, flushes into an
//            LDS node block of the tile; the block is written to a per-tile
//            partial buffer (no global atomics)
//   grid   : each owned node of a touched tile sums the partial blocks that
//            cover it, then v = p/m, v_update = v + dt (f + mb)/m, fixes
//   G2P    : one workgroup per active tile; LDS block of node velocities;
//            lanes = particles: gather, L, x/v update, F, stress, CFL, next
//            cell key; state is written in sorted order (the permutation is
//            applied for free)

#include "kmpm_kernels.h"

namespace kmpm {

// ---------------------------------------------------------------------------
// sort
// ---------------------------------------------------------------------------
__global__ void k_keys(KArgs a) {
  if (frozen(a)) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const int f = wpos(a);
  a.key[i] = particle_key(a, PF(a.pa, f, i), PF(a.pa, f + 1, i), PF(a.pa, f + 2, i));
}

// Resets the step's tile-list counters when no particle runs k_count (which
// resets them otherwise).  Per-tile counts are zero between sorts: k_scatter
// clears them once k_tiles has consumed them.
__global__ void k_sort_begin(KArgs a) {
  if (frozen(a)) return;
  if (threadIdx.x == 0) {
    a.sc->tile_count = 0;
    a.sc->dil_count = 0;
  }
}

// Particles per lane in k_count and k_scatter: a lane handles SORT_PPT
// particles a wavefront apart, so one wavefront has SORT_PPT loads or atomics
// in flight before it waits.  Wavefronts whose rows lie inside [0, np) take
// a path with no per-lane conditions around loads and stores, so that the
// compiler's wait counts stay exact and loads are not serialised behind the
// atomics and stores of earlier rows.  Cells per tile are a power of two.
constexpr int SORT_PPT = 8;
constexpr int SORT_BLOCK = 256;
constexpr int SORT_PER_WAVE = 64 * SORT_PPT;
constexpr int SORT_PER_BLOCK = SORT_BLOCK * SORT_PPT;

// Runs of equal tiles in one row of a wavefront: the run of this lane starts
// at lane `start` and ends before lane `end`; `head` marks the first lane.
struct TileRun {
  int start, end;
  bool head;
};

__device__ inline TileRun tile_run(int tile, bool valid, int lane) {
  const int prev = __shfl_up(tile, 1, 64);
  TileRun r;
  r.head = valid && (lane == 0 || prev != tile);
  const unsigned long long heads = __ballot(r.head);
  const unsigned long long live = __ballot(valid);
  const unsigned long long below = heads & ((2ull << lane) - 1ull);   // heads at lanes <= lane
  r.start = below ? 63 - __clzll(below) : 0;
  const unsigned long long above = lane < 63 ? (heads >> (lane + 1)) : 0ull;
  r.end = above ? lane + 1 + __ffsll(above) - 1 : __popcll(live);
  return r;
}

template <bool FULL>
__device__ inline void count_rows(const KArgs &a, int i0, int lane) {
  const int cshift = __ffs(a.tg.cpt) - 1;
  int key[SORT_PPT], tile[SORT_PPT];
#pragma unroll
  for (int k = 0; k < SORT_PPT; k++) {
    const int i = i0 + k * 64;
    key[k] = a.key[FULL ? i : min(i, a.np - 1)];
  }
#pragma unroll
  for (int k = 0; k < SORT_PPT; k++)
    tile[k] = FULL || i0 + k * 64 < a.np ? key[k] >> cshift : -1;
#pragma unroll
  for (int k = 0; k < SORT_PPT; k++) {
    const TileRun r = tile_run(tile[k], FULL || i0 + k * 64 < a.np, lane);
    if (r.head) atomicAdd(&a.tile_np[tile[k]], r.end - lane);
  }
  if (a.ninst) {   // batched contexts: the particle's stencil must stay in its instance
#pragma unroll
    for (int k = 0; k < SORT_PPT; k++) {
      const int i = i0 + k * 64;
      if (FULL || i < a.np) {
        int c[3];
        key_cell(a, key[k], c);
        check_instance(a, c, tile[k], IFLD(a.ia, PID, i));
      }
    }
  }
}

// Counts particles per tile.  Particles are stored in the previous step's
// sorted order, so equal tiles come in runs: one atomic per run of a tile
// inside a wavefront's row.
__global__ __launch_bounds__(SORT_BLOCK) void k_count(KArgs a) {
  if (frozen(a)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.sc->tile_count = 0;
    a.sc->dil_count = 0;
  }
  const int lane = threadIdx.x & 63;
  const int w0 = blockIdx.x * SORT_PER_BLOCK + (threadIdx.x >> 6) * SORT_PER_WAVE;
  if (w0 >= a.np) return;
  if (w0 + SORT_PER_WAVE <= a.np) count_rows<true>(a, w0 + lane, lane);
  else count_rows<false>(a, w0 + lane, lane);
}

template <bool FULL>
__device__ inline void scatter_rows(const KArgs &a, int i0, int lane) {
  const int cshift = __ffs(a.tg.cpt) - 1;
  int tile[SORT_PPT], base[SORT_PPT], start[SORT_PPT];
#pragma unroll
  for (int k = 0; k < SORT_PPT; k++) {
    const int i = i0 + k * 64;
    tile[k] = a.key[FULL ? i : min(i, a.np - 1)] >> cshift;
  }
#pragma unroll
  for (int k = 0; k < SORT_PPT; k++) {
    const TileRun r = tile_run(tile[k], FULL || i0 + k * 64 < a.np, lane);
    start[k] = r.start;
    base[k] = 0;
    if (r.head) base[k] = atomicAdd(&a.tile_cur[tile[k]], r.end - lane);
  }
#pragma unroll
  for (int k = 0; k < SORT_PPT; k++) {
    const int i = i0 + k * 64;
    const int pos = __shfl(base[k], start[k], 64) + (lane - start[k]);
    if (FULL || i < a.np) a.order_tmp[pos] = i;
  }
}

// Groups the slots by tile: each run of a tile in a wavefront's row takes a
// range of the tile's positions from its cursor.  The order inside a tile
// is settled by k_tile_sort.
__global__ __launch_bounds__(SORT_BLOCK) void k_scatter(KArgs a) {
  if (frozen(a)) return;
  // the per-tile counts are zero between sorts (k_tiles has read them all)
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t <= a.tg.ntiles; t += gridDim.x * blockDim.x)
    a.tile_np[t] = 0;
  const int lane = threadIdx.x & 63;
  const int w0 = blockIdx.x * SORT_PER_BLOCK + (threadIdx.x >> 6) * SORT_PER_WAVE;
  if (w0 >= a.np) return;
  if (w0 + SORT_PER_WAVE <= a.np) scatter_rows<true>(a, w0 + lane, lane);
  else scatter_rows<false>(a, w0 + lane, lane);
}

// Tile offsets: an exclusive scan of the per-tile counts (tiles + 1 values,
// the last zero) in two passes with no waits between workgroups.
// k_scan_sums writes the total of each chunk of SCAN_BLOCK counts; k_tiles
// adds the totals of the chunks before its own to a scan of its chunk.
constexpr int SCAN_BLOCK = 256;

__device__ inline int wave_sum_i(int v) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Sum of v over the block (every thread gets it); ws holds 4 ints.
__device__ inline int block_sum_i(int v, int *ws) {
  v = wave_sum_i(v);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  const int s = ws[0] + ws[1] + ws[2] + ws[3];
  __syncthreads();
  return s;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(const int *cnt, int n, int *part) {
  __shared__ int ws[4];
  const int i = blockIdx.x * SCAN_BLOCK + (int)threadIdx.x;
  const int s = block_sum_i(i < n ? cnt[i] : 0, ws);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// One thread per tile: its first position (the scan), then the active tiles
// (particles in the tile) and the tiles whose owned nodes lie in the node
// block of an active tile: the active tile itself and its neighbours at -1
// (and +1 for cubic) per axis, from the counts.  Sets the scatter cursors.
// The counts are cleared by k_scatter, after every tile has read its
// neighbours'.
__global__ __launch_bounds__(SCAN_BLOCK) void k_tiles(KArgs a, const int *part) {
  if (frozen(a)) return;
  __shared__ int ws[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int t = blockIdx.x * SCAN_BLOCK + tid;
  const TileGeom &g = a.tg;
  int o = 0;
  for (int q = tid; q < (int)blockIdx.x; q += SCAN_BLOCK) o += part[q];
  const int base = block_sum_i(o, ws);
  const int cnt = t <= g.ntiles ? a.tile_np[t] : 0;
  int x = cnt;   // inclusive prefix inside the wave, then the waves before
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) ws[wave] = x;
  __syncthreads();
  int p0 = base + x - cnt;
  for (int w = 0; w < wave; w++) p0 += ws[w];
  if (t <= g.ntiles) a.tile_p0[t] = p0;
  const bool in = t < g.ntiles;
  bool active = false, hit = false;
  if (in) {
    a.tile_cur[t] = p0;
    active = cnt > 0;
    const int tz = t % g.nt[2], ty = (t / g.nt[2]) % g.nt[1], tx = t / (g.nt[2] * g.nt[1]);
    // all neighbours' counts in flight at once (no early exit)
    const int up = a.S1 == 4 ? 1 : 0;
    const int zlo = a.dim == 3 ? -1 : 0, zhi = a.dim == 3 ? up : 0;
#pragma unroll
    for (int dx = -1; dx <= 1; dx++)
#pragma unroll
      for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dz = -1; dz <= 1; dz++) {
          const int ux = tx + dx, uy = ty + dy, uz = tz + dz;
          const bool use = dx <= up && dy <= up && dz >= zlo && dz <= zhi && ux >= 0 &&
                           uy >= 0 && uz >= 0 && ux < g.nt[0] && uy < g.nt[1] && uz < g.nt[2];
          const int u = use ? (ux * g.nt[1] + uy) * g.nt[2] + uz : t;
          hit |= a.tile_np[u] > 0;
        }
  }
  // one atomic per wavefront appends to both lists (the counts are the low
  // and high halves of one 64-bit word, neither exceeds ntiles)
  const unsigned long long ma = __ballot(active), mh = __ballot(hit);
  const int leader = __ffsll(ma | mh | 1ull) - 1;
  unsigned long long lb = 0;
  if (lane == leader && (ma | mh))
    lb = atomicAdd((unsigned long long *)&a.sc->tile_count,
                   (unsigned long long)__popcll(ma) | ((unsigned long long)__popcll(mh) << 32));
  lb = __shfl(lb, leader, 64);
  const unsigned long long below = (1ull << lane) - 1ull;
  const int aslot = (int)(lb & 0xffffffffull) + __popcll(ma & below);
  const int hslot = (int)(lb >> 32) + __popcll(mh & below);
  if (active) a.tile_list[aslot] = t;
  if (in) a.tile_slot[t] = active ? aslot : -1;
  if (hit) a.dil_list[hslot] = t;
}

// Orders the slots of each active tile by cell, ties by slot: a stable
// counting sort, since the slots are the previous step's sorted order (the
// upload order, ascending tag, before the first step).  A counting sort over
// the tile's cells in LDS, then each slot's rank among the smaller slots of
// its cell.  One wavefront per tile, so that many tiles are in flight per CU.
// A tile of up to TILE_SORT_CAP particles is read once into registers
// (TILE_SORT_KPL per lane) and its slots placed by cell in LDS; a larger one
// places them in its own range of sort_tmp and finds a position's cell by
// binary search of the offsets.  The next tile's range is fetched while the
// current one is ordered.  Dynamic LDS: tile_sort_lds_bytes(cpt).
__global__ __launch_bounds__(64) void k_tile_sort(KArgs a) {
  if (frozen(a)) return;
  extern __shared__ int lds[];
  const int cpt = a.tg.cpt, cmask = cpt - 1;
  int *off = lds;                       // [cpt + 1] first position of each cell
  int *cur = off + cpt + 1;             // [cpt] counts, then placement cursors
  int *s_slot = cur + cpt;              // [TILE_SORT_CAP] slots placed by cell
  const int lane = threadIdx.x;
  const int per = (cpt + 63) >> 6;      // cells per lane in the scan
  const int ntile = a.sc->tile_count;
  int li = blockIdx.x;
  int p0 = 0, p1 = 0;
  if (li < ntile) {
    const int t = a.tile_list[li];
    p0 = a.tile_p0[t];
    p1 = a.tile_p0[t + 1];
  }
  for (; li < ntile; li += gridDim.x) {
    const int n = p1 - p0;
    const int *in = a.order_tmp + p0;
    const bool inreg = n <= TILE_SORT_CAP;
    int sl[TILE_SORT_KPL], ce[TILE_SORT_KPL];
#pragma unroll
    for (int k = 0; k < TILE_SORT_KPL; k++) {
      const int j = k * 64 + lane;
      sl[k] = inreg && j < n ? in[j] : -1;
    }
#pragma unroll
    for (int k = 0; k < TILE_SORT_KPL; k++) ce[k] = sl[k] >= 0 ? a.key[sl[k]] & cmask : -1;
    // the next tile's range, in flight during this tile's ordering
    const int nli = li + gridDim.x;
    const int nt = nli < ntile ? a.tile_list[nli] : 0;
    __syncthreads();   // the previous tile is done with LDS
    for (int c = lane; c < cpt; c += 64) cur[c] = 0;
    __syncthreads();
    if (inreg) {
#pragma unroll
      for (int k = 0; k < TILE_SORT_KPL; k++)
        if (ce[k] >= 0) atomicAdd(&cur[ce[k]], 1);
    } else {
      for (int j = lane; j < n; j += 64) atomicAdd(&cur[a.key[in[j]] & cmask], 1);
    }
    __syncthreads();
    {   // exclusive scan of the counts, `per` consecutive cells per lane
      const int b = lane * per;
      int sum = 0;
      for (int k = 0; k < per; k++) sum += b + k < cpt ? cur[b + k] : 0;
      int x = sum;
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
      }
      int ex = x - sum;
      for (int k = 0; k < per; k++)
        if (b + k < cpt) {
          const int v = cur[b + k];
          off[b + k] = ex;
          cur[b + k] = ex;
          ex += v;
        }
      if (lane == 0) off[cpt] = n;
    }
    __syncthreads();
    int np0 = 0, np1 = 0;
    if (nli < ntile) {
      np0 = a.tile_p0[nt];
      np1 = a.tile_p0[nt + 1];
    }
    if (inreg) {
#pragma unroll
      for (int k = 0; k < TILE_SORT_KPL; k++)
        if (ce[k] >= 0) s_slot[atomicAdd(&cur[ce[k]], 1)] = sl[k];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < TILE_SORT_KPL; k++)
        if (ce[k] >= 0) {
          const int s = off[ce[k]], q1 = off[ce[k] + 1];
          int r = 0;
          for (int q = s; q < q1; q++) r += s_slot[q] < sl[k] ? 1 : 0;
          a.order[p0 + s + r] = sl[k];
        }
    } else {
      int *g = a.sort_tmp + p0;
      for (int j = lane; j < n; j += 64) {
        const int v = in[j];
        g[atomicAdd(&cur[a.key[v] & cmask], 1)] = v;
      }
      __threadfence_block();
      __syncthreads();
      for (int j = lane; j < n; j += 64) {
        int lo = 0, hi = cpt;   // the last cell with off <= j: off[lo] <= j < off[lo + 1]
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (off[mid] <= j) lo = mid;
          else hi = mid;
        }
        const int me = g[j];
        const int s = off[lo], q1 = off[lo + 1];
        int r = 0;
        for (int q = s; q < q1; q++) r += g[q] < me ? 1 : 0;
        a.order[p0 + s + r] = me;
      }
    }
    p0 = np0;
    p1 = np1;
  }
}

size_t tile_sort_lds_bytes(int cpt) {
  return (size_t)(2 * cpt + 1 + TILE_SORT_CAP) * sizeof(int);
}

// update_time + adjust_dt (update.cpp:201-206, ulmpm.cpp:525-551)
// Batched contexts: every instance advances its own time and takes the dt
// of its own CFL condition (one thread per instance).
__device__ inline void adjust_dt_instances(const KArgs &a) {
  for (int i = threadIdx.x; i < a.ninst; i += blockDim.x) {
    InstScalars &s = a.inst[i];
    s.time += s.dt;
    if (!a.dt_constant) {
      const double maxc = __longlong_as_double((long long)s.max_c_bits);
      double dtcfl = a.h / maxc;
      dtcfl = 1.0e22 < dtcfl ? 1.0e22 : dtcfl;
      if (dtcfl == 0 || isnan(dtcfl)) set_error(a.sc, KMPM_ERR_DTCFL, -1);
      s.dt = dtcfl * a.dt_factor;
    }
    s.max_c_bits = 0;
    s.min_ratio_bits = (unsigned long long)__double_as_longlong(1.0);
  }
  if (threadIdx.x == 0) {
    a.sc->step += 1;
    for (int q = 0; q < MAX_FIX * 3; q++) a.sc->fv_force[a.sc->step & 1][q / 3][q % 3] = 0.0;
  }
}

// The end of a step: time and dt (update.cpp:201-206, ulmpm.cpp:525-551).  A
// step that raised a sticky error freezes the context (see frozen()); a
// frozen context's time, dt and step count stay as they are.
__global__ void k_adjust_dt(KArgs a) {
  if (blockIdx.x != 0 || frozen(a)) return;
  if (a.ninst) {
    adjust_dt_instances(a);
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd(&a.sc->error, 0) != 0) a.sc->frozen = 1;
    return;
  }
  if (threadIdx.x != 0) return;
  DevScalars *sc = a.sc;
  sc->time += sc->dt;
  if (!a.dt_constant) {
    const double maxc = __longlong_as_double((long long)sc->max_c_bits);
    // TLMPM scales by the smallest |Re eig(F)| (solid.cpp:1425); ULMPM by 1
    const double ratio = a.tl ? __longlong_as_double((long long)sc->min_ratio_bits) : 1.0;
    double dtcfl = a.h * ratio / maxc;
    dtcfl = 1.0e22 < dtcfl ? 1.0e22 : dtcfl;
    if ((dtcfl == 0 || isnan(dtcfl)) && sc->error == 0) {
      sc->error = KMPM_ERR_DTCFL;
      sc->err_pid = -1;
    }
    sc->dt = dtcfl * a.dt_factor;
  }
  sc->max_c_bits = 0;
  sc->min_ratio_bits = (unsigned long long)__double_as_longlong(1.0);
  sc->step += 1;
  for (int q = 0; q < MAX_FIX * 3; q++) sc->fv_force[sc->step & 1][q / 3][q % 3] = 0.0;
  if (sc->error != 0) sc->frozen = 1;
}

// ---------------------------------------------------------------------------
// observables and debug
// ---------------------------------------------------------------------------
template <int PREC>
__global__ void k_energy(KArgs a, const uint32_t *pmask, uint32_t bit, double *partial) {
  __shared__ double sk[256], ss[256];
  double ek = 0, es = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.np; i += gridDim.x * blockDim.x) {
    const int pid = IFLD(a.ia, PID, i);
    if (!(pmask[pid] & bit)) continue;
    const double v0 = LDA(VX, i), v1 = LDA(VY, i), v2 = LDA(VZ, i);
    const double nv = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    ek += 0.5 * LDA(MASS, i) * (nv * nv);
    // sigma : strain_el over the full tensor, row-major order
    const double s00 = LDA(SXX, i), s11 = LDA(SYY, i), s22 = LDA(SZZ, i);
    const double s01 = LDA(SXY, i), s02 = LDA(SXZ, i), s12 = LDA(SYZ, i);
    const double e00 = LDA(EXX, i), e11 = LDA(EYY, i), e22 = LDA(EZZ, i);
    const double e01 = LDA(EXY, i), e02 = LDA(EXZ, i), e12 = LDA(EYZ, i);
    double t = s00 * e00;
    t += s01 * e01;
    t += s02 * e02;
    t += s01 * e01;
    t += s11 * e11;
    t += s12 * e12;
    t += s02 * e02;
    t += s12 * e12;
    t += s22 * e22;
    es += 0.5 * LDA(VOL, i) * t;
  }
  sk[threadIdx.x] = ek;
  ss[threadIdx.x] = es;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sk[threadIdx.x] += sk[threadIdx.x + o];
      ss[threadIdx.x] += ss[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sk[0];
    partial[2 * blockIdx.x + 1] = ss[0];
  }
}

__global__ void k_energy_final(const double *partial, int nb, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double ek = 0, es = 0;
  for (int b = 0; b < nb; b++) {
    ek += partial[2 * b];
    es += partial[2 * b + 1];
  }
  out[0] = ek;
  out[1] = es;
}

// Dense stencil of the current positions (per pid): node tag or -1 when the
// node does not exist or the pair is dropped, w, grad w.
template <int S1>
__global__ void k_debug_weights(KArgs a, int64_t *tag_out, double *w_out, double *wd_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const int S = a.dim == 3 ? S1 * S1 * S1 : S1 * S1;
  const int off = S1 == 4 ? 1 : 0;
  const int pid = IFLD(a.ia, PID, i);
  double w[3][S1], dw[3][S1];
  int base[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++)
    for (int q = 0; q < S1; q++) { w[d][q] = 1; dw[d][q] = 0; }
  for (int d = 0; d < a.dim; d++) {
    const double x = PF(a.pa, wpos(a) + d, i);
    const int c = cell_of_clamped(x, a.lo[d], a.inv_h, a.n[d]);
    base[d] = c - off;
    weights_1d<S1>(x, a.lo[d], a.h, a.inv_h, a.n[d], c - off, w[d], dw[d]);
  }
  int s = 0;
  for (int ia = 0; ia < S1; ia++)
    for (int ib = 0; ib < S1; ib++)
      for (int ic = 0; ic < (a.dim == 3 ? S1 : 1); ic++, s++) {
        const int64_t o = (int64_t)pid * S + s;
        double wt, g0, g1, g2;
        if (a.dim == 3) {
          const double wxy = w[0][ia] * w[1][ib];
          wt = wxy * w[2][ic];
          g0 = dw[0][ia] * w[1][ib] * w[2][ic];
          g1 = w[0][ia] * dw[1][ib] * w[2][ic];
          g2 = wxy * dw[2][ic];
        } else {
          wt = w[0][ia] * w[1][ib];
          g0 = dw[0][ia] * w[1][ib];
          g1 = w[0][ia] * dw[1][ib];
          g2 = 0;
        }
        const int ni = base[0] + ia, nj = base[1] + ib, nk = a.dim == 3 ? base[2] + ic : 0;
        const bool ok = wt != 0 && ni >= 0 && nj >= 0 && nk >= 0 && ni < a.n[0] &&
                        nj < a.n[1] && nk < a.n[2];
        tag_out[o] = ok ? ((int64_t)ni * a.n[1] + nj) * a.n[2] + nk : -1;
        w_out[o] = ok ? wt : 0;
        wd_out[3 * o] = ok ? g0 : 0;
        wd_out[3 * o + 1] = ok ? g1 : 0;
        wd_out[3 * o + 2] = ok ? g2 : 0;
      }
}

// Reference particle-cell indices (ulmpm.cpp:163-165 linear, :206-208 splines), by pid.
__global__ void k_cells(KArgs a, int *ijk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const int pid = IFLD(a.ia, PID, i);
  for (int d = 0; d < 3; d++) {
    int v = 0;
    if (d < a.dim) {
      const double x = PF(a.pa, wpos(a) + d, i);
      if (a.S1 == 2) v = (int)((x - a.lo[d]) * a.inv_h);
      else v = (int)((x - a.lo[d]) * a.inv_h - 1);
    }
    ijk[3 * (int64_t)pid + d] = v;
  }
}

// ---------------------------------------------------------------------------
// tile configurations and launchers
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void tile_shape(int dim, int S1, int variant, int T[3]) {
#define KMPM_SHAPE(D, S, V)                                               \
  if (dim == D && S1 == S && variant == V) {                              \
    T[0] = Tiles<D, S, V>::TX;                                            \
    T[1] = Tiles<D, S, V>::TY;                                            \
    T[2] = D == 3 ? Tiles<D, S, V>::TZ : 1;                               \
    return;                                                               \
  }
  KMPM_SHAPE(3, 4, 0) KMPM_SHAPE(3, 4, 1) KMPM_SHAPE(3, 4, 2)
  KMPM_SHAPE(3, 2, 0) KMPM_SHAPE(3, 2, 1) KMPM_SHAPE(3, 2, 2)
  KMPM_SHAPE(2, 4, 0) KMPM_SHAPE(2, 4, 1) KMPM_SHAPE(2, 4, 2)
  KMPM_SHAPE(2, 2, 0) KMPM_SHAPE(2, 2, 1) KMPM_SHAPE(2, 2, 2)
#undef KMPM_SHAPE
  T[0] = T[1] = T[2] = 0;
}

// The context's (dimension, shape function, tile variant) selects the
// TileLaunch instantiation.
#define TILE_CALL(c, CALL)                                                          \
  do {                                                                              \
    switch ((c)->a.dim * 100 + (c)->a.S1 * 10 + (c)->tile_variant) {                \
      case 340: return TileLaunch<3, 4, 0>::CALL;                                   \
      case 341: return TileLaunch<3, 4, 1>::CALL;                                   \
      case 342: return TileLaunch<3, 4, 2>::CALL;                                   \
      case 320: return TileLaunch<3, 2, 0>::CALL;                                   \
      case 321: return TileLaunch<3, 2, 1>::CALL;                                   \
      case 322: return TileLaunch<3, 2, 2>::CALL;                                   \
      case 240: return TileLaunch<2, 4, 0>::CALL;                                   \
      case 241: return TileLaunch<2, 4, 1>::CALL;                                   \
      case 242: return TileLaunch<2, 4, 2>::CALL;                                   \
      case 220: return TileLaunch<2, 2, 0>::CALL;                                   \
      case 221: return TileLaunch<2, 2, 1>::CALL;                                   \
      case 222: return TileLaunch<2, 2, 2>::CALL;                                   \
    }                                                                               \
    return hipErrorInvalidValue;                                                    \
  } while (0)

hipError_t launch_p2g(kmpm_ctx *c, int remap) { TILE_CALL(c, p2g(c, remap)); }
// The expression-valued fixes' values of this step into sc->fv_last (one
// block of 64 threads, launched before every grid pass of a context with
// such fixes, so the interpreter's registers stay out of the grid kernels).
// Expressions see the time at the start of the step and the script's dt: 0
// in the first step unless set_dt (input.cpp:55, update.cpp:69-76), the
// step's dt after that (ulmpm.cpp:550).
__global__ __launch_bounds__(64) void k_fix_values(KArgs a) {
  const int tid = threadIdx.x;
  if (frozen(a) || tid >= a.nfv * 6) return;
  const int fi = tid / 6, w = tid % 6;
  if (!a.fv[fi].expr) return;
  const double dtv = (a.sc->step == 0 && !a.dt_constant) ? 0.0 : a.sc->dt;
  a.sc->fv_last[fi][w] = expr_eval(a.fexpr[6 * fi + w], a.sc->time, dtv);
}

static bool has_expr_fix(const kmpm_ctx *c) {
  for (int fi = 0; fi < c->a.nfv; fi++)
    if (c->a.fv[fi].expr) return true;
  return false;
}

hipError_t launch_fix_values(kmpm_ctx *c) {
  if (!has_expr_fix(c)) return hipSuccess;
  hipLaunchKernelGGL(k_fix_values, dim3(1), dim3(64), 0, c->stream, c->a);
  return hipGetLastError();
}

static hipError_t tile_grid_update(kmpm_ctx *c, int mode) { TILE_CALL(c, grid_update(c, mode)); }

hipError_t launch_grid_update(kmpm_ctx *c, int mode) {
  const hipError_t e = launch_fix_values(c);
  if (e != hipSuccess) return e;
  return tile_grid_update(c, mode);
}
hipError_t launch_grid_clean(kmpm_ctx *c) { TILE_CALL(c, grid_clean(c)); }
hipError_t launch_p2g_thermal(kmpm_ctx *c, int remap) { TILE_CALL(c, p2g_thermal(c, remap)); }
hipError_t launch_grid_thermal(kmpm_ctx *c, int mode) { TILE_CALL(c, grid_thermal(c, mode)); }
hipError_t launch_g2p_usl(kmpm_ctx *c) { TILE_CALL(c, g2p_usl(c)); }
hipError_t launch_g2p_musl_a(kmpm_ctx *c) { TILE_CALL(c, g2p_musl_a(c)); }
hipError_t launch_stress_musl_b(kmpm_ctx *c) { TILE_CALL(c, stress_musl_b(c)); }

hipError_t launch_keys(kmpm_ctx *c) {
  if (c->a.np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_keys, dim3(nblk(c->a.np, 256)), dim3(256), 0, c->stream, c->a);
  return hipGetLastError();
}

size_t scan_temp_bytes(int64_t n) {
  return (size_t)((n + SCAN_BLOCK - 1) / SCAN_BLOCK) * sizeof(int);
}

hipError_t launch_sort(kmpm_ctx *c) {
  KArgs &a = c->a;
  const int64_t nt = a.tg.ntiles;
  if (a.np > 0)
    hipLaunchKernelGGL(k_count, dim3(nblk(a.np, SORT_PER_BLOCK)), dim3(SORT_BLOCK), 0, c->stream,
                       a);
  else
    hipLaunchKernelGGL(k_sort_begin, dim3(1), dim3(64), 0, c->stream, a);
  {
    const int nb = (int)nblk(nt + 1, SCAN_BLOCK);
    int *part = static_cast<int *>(c->d_scan_tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(nb), dim3(SCAN_BLOCK), 0, c->stream, a.tile_np,
                       (int)(nt + 1), part);
    hipLaunchKernelGGL(k_tiles, dim3(nb), dim3(SCAN_BLOCK), 0, c->stream, a, (const int *)part);
  }
  if (a.np > 0) {
    hipLaunchKernelGGL(k_scatter, dim3(nblk(a.np, SORT_PER_BLOCK)), dim3(SORT_BLOCK), 0,
                       c->stream, a);
    hipLaunchKernelGGL(k_tile_sort, dim3(persistent_grid(c, 32)), dim3(64),
                       tile_sort_lds_bytes(a.tg.cpt), c->stream, a);
  }
  return hipGetLastError();
}

hipError_t launch_adjust_dt(kmpm_ctx *c) {
  hipLaunchKernelGGL(k_adjust_dt, dim3(1), dim3(256), 0, c->stream, c->a);
  return hipGetLastError();
}

hipError_t launch_energy(kmpm_ctx *c, uint32_t bit, double *d_out2) {
  const int nb = ENERGY_BLOCKS;
  PREC_DISPATCH(c, hipLaunchKernelGGL(k_energy<PREC>, dim3(nb), dim3(256), 0, c->stream, c->a,
                                      c->d_pmask, bit, d_out2 + 2));
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(64), 0, c->stream, d_out2 + 2, nb, d_out2);
  return hipGetLastError();
}

hipError_t launch_debug_weights(kmpm_ctx *c, int64_t *d_tag, double *d_w, double *d_wd) {
  if (c->a.np == 0) return hipSuccess;
  if (c->a.S1 == 4)
    hipLaunchKernelGGL(k_debug_weights<4>, dim3(nblk(c->a.np, 128)), dim3(128), 0, c->stream,
                       c->a, d_tag, d_w, d_wd);
  else
    hipLaunchKernelGGL(k_debug_weights<2>, dim3(nblk(c->a.np, 128)), dim3(128), 0, c->stream,
                       c->a, d_tag, d_w, d_wd);
  return hipGetLastError();
}

hipError_t launch_cells(kmpm_ctx *c, int *d_ijk) {
  if (c->a.np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cells, dim3(nblk(c->a.np, 256)), dim3(256), 0, c->stream, c->a, d_ijk);
  return hipGetLastError();
}

__global__ void k_iota(int *d, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = i;
}

hipError_t launch_iota(int *d, int n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3(nblk(n, 256)), dim3(256), 0, stream, d, n);
  return hipGetLastError();
}

// Streaming copy used to calibrate hardware byte counters.
__global__ void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

hipError_t copy_probe(int64_t bytes, int reps, double *gbps) {
  const int64_t n = bytes / (int64_t)sizeof(double2);
  double2 *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, n * sizeof(double2));
  if (e == hipSuccess) e = hipMalloc(&b, n * sizeof(double2));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) {
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int grid = cus * 8;
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, a, b, n);   // warm-up
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, a, b, n);
    (void)hipEventRecord(e1, s);
    e = hipStreamSynchronize(s);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) *gbps = 2.0 * (double)n * sizeof(double2) * reps / (ms * 1e-3) / 1e9;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return e;
}

}  // namespace kmpm
