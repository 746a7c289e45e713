// Domain construction on the GPU from a flattened Mesh (Domain::from_mesh,
// domain.rs:69-159, with Mesh::set_edge_activation, mesh.rs:1180-1215).
//
// The sequential algorithm generates every Elem's BasisSpecs, numbers the
// elem-type DoFs of the leaves in Elem order, then the matched edge-type DoFs
// Edge by Edge, and appends each DoF's specs to the lists of their Elems as it
// goes. Each of these orders is a prefix sum over counts that depend on one
// Elem or one Edge only, and the counts have closed forms:
//  - the elem-type specs of an Elem with orders (ni, nj) are the U specs with
//    j >= 2 and the V specs with i >= 2 (basis_spec.rs:41-75):
//    ni (nj - 1) + (ni - 1) nj of them, U first, i outer, j inner;
//  - the U/V specs of an Elem on its local Edge L are (U, i, L), i < ni, for
//    L = 0, 1 and (V, L - 2, j), j < nj, for L = 2, 3, in generation order;
//  - on an active Edge with pair A < B (local Edges La, Lb), the first-match
//    pairing of matches_with_edge (basis_spec.rs:80-110) pairs A's k-th spec
//    with B's k-th when La + Lb = 1 (U) or 5 (V), for k below the smaller of
//    the two counts, and pairs nothing otherwise (no two specs of one Elem on
//    one Edge match: their local indices would have to differ).
// So the Domain comes out of: an activation and pair count per Edge, a count
// per Elem, three prefix sums, and one pass writing every Elem's list.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "domain_build.hpp"

namespace fem2d {
namespace build {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kItems = 8;                    // per thread in the prefix sums
constexpr int kTile = kBlock * kItems;       // items per prefix-sum tile
constexpr long long kGroupCountBlocks = 512;  // blocks of the row-group count (one atomic each)

__device__ __forceinline__ bool has_pair(int2 s) { return s.x >= 0 && s.y >= 0; }

// local index (0..3) of Edge e in an Elem's Edge list, -1 if absent
__device__ __forceinline__ int local_edge(int4 ed, int e) {
  return ed.x == e ? 0 : ed.y == e ? 1 : ed.z == e ? 2 : ed.w == e ? 3 : -1;
}

// Edge::set_activation over the recursion of Mesh::set_edge_activation: the
// recursion starts at every base Edge off the boundary and descends into an
// Edge's children only while the Edge has a candidate pair (an Elem on each
// side); an Edge it reaches with a candidate pair stays active unless both of
// its children have one. One thread per Edge walks its ancestors.
__global__ void __launch_bounds__(kBlock) k_edges(Args a) {
  const int e = int(blockIdx.x) * kBlock + int(threadIdx.x);
  if (e >= a.n_edges) return;
  const int2 s = a.side[e];
  const bool cand = has_pair(s);
  bool reached = true;
  int x = e;
  for (int steps = 0;; ++steps) {
    const int p = a.edge_parent[x];
    if (p < 0) break;
    if (steps == kMaxEdgeDepth) {
      atomicOr(&a.status->flags, kFlagTooDeep);
      reached = false;
      break;
    }
    if (!has_pair(a.side[p])) {
      reached = false;
      break;
    }
    x = p;
  }
  reached = reached && a.open_base[x] != 0;
  bool active = false;
  if (reached) {
    if (!cand) {
      if (a.edge_parent[e] < 0) atomicOr(&a.status->flags, kFlagNoPair);
    } else {
      const int2 c = a.child[e];
      if (c.x >= 0) {
        const bool c0 = has_pair(a.side[c.x]), c1 = has_pair(a.side[c.y]);
        active = !(c0 && c1);
        if (c0 != c1) atomicOr(&a.status->flags, kFlagInconsistent);
      } else {
        active = true;
      }
    }
  }
  int2 pr = make_int2(-1, -1);
  int np = 0;
  if (active) {
    const int A = min(s.x, s.y), B = max(s.x, s.y);
    pr = make_int2(A, B);
    if (A != B) {
      const int la = local_edge(a.elem_edges[A], e), lb = local_edge(a.elem_edges[B], e);
      const unsigned ma = a.elem_meta[A], mb = a.elem_meta[B];
      if (la >= 0 && lb >= 0) {
        if (la < 2 && lb < 2 && la + lb == 1)
          np = int(min(ma & 255u, mb & 255u));
        else if (la >= 2 && lb >= 2 && la + lb == 5)
          np = int(min((ma >> 8) & 255u, (mb >> 8) & 255u));
      }
    }
  }
  a.act[e] = pr;
  a.n_pairs[e] = np;
}

// One thread per Elem: its elem-type DoFs (leaves only) and the length of its
// list (those, plus its spec of every pair on each of its Edges)
__global__ void __launch_bounds__(kBlock) k_elems(Args a) {
  const int x = int(blockIdx.x) * kBlock + int(threadIdx.x);
  if (x >= a.n_elems) return;
  const unsigned m = a.elem_meta[x];
  const int ni = int(m & 255u), nj = int((m >> 8) & 255u);
  const int n_elem = (m >> 16) & 1u ? ni * (nj - 1) + (ni - 1) * nj : 0;
  const int4 ed = a.elem_edges[x];
  int len = n_elem;
  const int es[4] = {ed.x, ed.y, ed.z, ed.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int2 p = a.act[es[k]];
    if (p.x != p.y && (p.x == x || p.y == x)) len += a.n_pairs[es[k]];
  }
  a.n_elem[x] = n_elem;
  a.list_len[x] = len;
}

// ---------------------------------------------------------------------------
// Three exclusive prefix sums in one set of launches: the elem-type DoFs per
// Elem, the pairs per Edge and the list lengths per Elem. Tiles of kTile items
// (at least one per array); each array's output has one entry more, its total.
// ---------------------------------------------------------------------------

struct ScanSeg {
  const int* in;
  long long* out;
  long long n;
  long long first_tile;
};

__device__ __forceinline__ ScanSeg scan_seg(const Args& a, int which) {
  const long long t0 = (a.n_elems + kTile - 1) / kTile > 0 ? (a.n_elems + kTile - 1) / kTile : 1;
  const long long t1 = (a.n_edges + kTile - 1) / kTile > 0 ? (a.n_edges + kTile - 1) / kTile : 1;
  if (which == 0) return {a.n_elem, a.elem_base, a.n_elems, 0};
  if (which == 1) return {a.n_pairs, a.edge_base, a.n_edges, t0};
  return {a.list_len, a.bs_off, a.n_elems, t0 + t1};
}

__device__ __forceinline__ int seg_of_tile(const Args& a, long long b) {
  const long long t0 = (a.n_elems + kTile - 1) / kTile > 0 ? (a.n_elems + kTile - 1) / kTile : 1;
  const long long t1 = (a.n_edges + kTile - 1) / kTile > 0 ? (a.n_edges + kTile - 1) / kTile : 1;
  return b < t0 ? 0 : b < t0 + t1 ? 1 : 2;
}

__device__ __forceinline__ long long wave_incl_scan(long long x, int lane) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const long long y = __shfl_up(x, d, kWave);
    if (lane >= d) x += y;
  }
  return x;
}

// exclusive block-wide prefix sum (kBlock threads, every thread calls it)
__device__ __forceinline__ long long block_excl_scan(long long x, long long* wave_sum, long long* total) {
  const int lane = int(threadIdx.x) % kWave, wid = int(threadIdx.x) / kWave;
  const long long incl = wave_incl_scan(x, lane);
  if (lane == kWave - 1) wave_sum[wid] = incl;
  __syncthreads();
  long long before = 0, all = 0;
#pragma unroll
  for (int i = 0; i < kWavesPerBlock; ++i) {
    before += i < wid ? wave_sum[i] : 0;
    all += wave_sum[i];
  }
  __syncthreads();
  *total = all;
  return before + incl - x;
}

// the kItems items of this thread in tile `tile` of a segment (thread-contiguous)
__device__ __forceinline__ void load_items(const ScanSeg& s, long long tile, int (&v)[kItems]) {
  const long long i0 = tile * kTile + (long long)threadIdx.x * kItems;
#pragma unroll
  for (int k = 0; k < kItems; ++k) v[k] = i0 + k < s.n ? s.in[i0 + k] : 0;
}

__global__ void __launch_bounds__(kBlock) k_scan_reduce(Args a) {
  __shared__ long long wave_sum[kWavesPerBlock];
  const long long b = blockIdx.x;
  const ScanSeg s = scan_seg(a, seg_of_tile(a, b));
  int v[kItems];
  load_items(s, b - s.first_tile, v);
  long long sum = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) sum += v[k];
  long long total = 0;
  (void)block_excl_scan(sum, wave_sum, &total);
  if (threadIdx.x == 0) a.tile_sums[b] = total;
}

// one block: each segment's tile sums -> tile offsets; the totals
__global__ void __launch_bounds__(kBlock) k_scan_top(Args a) {
  __shared__ long long wave_sum[kWavesPerBlock];
  for (int which = 0; which < 3; ++which) {
    const ScanSeg s = scan_seg(a, which);
    const long long n_tiles = (s.n + kTile - 1) / kTile > 0 ? (s.n + kTile - 1) / kTile : 1;
    long long carry = 0;
    for (long long c = 0; c < n_tiles; c += kBlock) {
      const long long t = c + threadIdx.x;
      const long long x = t < n_tiles ? a.tile_sums[s.first_tile + t] : 0;
      long long total = 0;
      const long long ex = block_excl_scan(x, wave_sum, &total);
      if (t < n_tiles) a.tile_sums[s.first_tile + t] = carry + ex;
      carry += total;
    }
    if (threadIdx.x == 0) {
      s.out[s.n] = carry;
      if (which == 0) a.status->elem_dofs = carry;
      if (which == 1) a.status->edge_dofs = carry;
      if (which == 2) a.status->specs = carry;
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(Args a) {
  __shared__ long long wave_sum[kWavesPerBlock];
  const long long b = blockIdx.x;
  const ScanSeg s = scan_seg(a, seg_of_tile(a, b));
  const long long tile = b - s.first_tile;
  int v[kItems];
  load_items(s, tile, v);
  long long sum = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) sum += v[k];
  long long total = 0;
  long long run = a.tile_sums[b] + block_excl_scan(sum, wave_sum, &total);
  const long long i0 = tile * kTile + (long long)threadIdx.x * kItems;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    if (i0 + k < s.n) s.out[i0 + k] = run;
    run += v[k];
  }
}

// ---------------------------------------------------------------------------
// The lists: one wavefront per Elem, its lanes over the list's positions. The
// list is the Elem's elem-type specs (leaves; their DoFs in generation order),
// then for each of its Edges in ascending id its spec of each pair on it.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void order2(int& ia, int& la, int& ib, int& lb) {
  if (ib < ia) {
    const int t = ia, u = la;
    ia = ib;
    la = lb;
    ib = t;
    lb = u;
  }
}

template <class T>
__device__ __forceinline__ T pick4(int k, T v0, T v1, T v2, T v3) {
  return k == 0 ? v0 : k == 1 ? v1 : k == 2 ? v2 : v3;
}

__global__ void __launch_bounds__(kBlock) k_lists(Args a) {
  const int lane = int(threadIdx.x) % kWave;
  const int x = int(blockIdx.x) * kWavesPerBlock + int(threadIdx.x) / kWave;
  if (x >= a.n_elems) return;
  const unsigned m = a.elem_meta[x];
  const int ni = int(m & 255u), nj = int((m >> 8) & 255u);
  const int n_elem = (m >> 16) & 1u ? ni * (nj - 1) + (ni - 1) * nj : 0;
  const int4 ed = a.elem_edges[x];
  // the Elem's Edges in ascending id, each with its local index
  int i0 = ed.x, l0 = 0, i1 = ed.y, l1 = 1, i2 = ed.z, l2 = 2, i3 = ed.w, l3 = 3;
  order2(i0, l0, i1, l1);
  order2(i2, l2, i3, l3);
  order2(i0, l0, i2, l2);
  order2(i1, l1, i3, l3);
  order2(i1, l1, i2, l2);
  int cnt[4], side[4];
  long long base[4];
  const int ids[4] = {i0, i1, i2, i3};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int2 p = a.act[ids[k]];
    const bool mine = p.x != p.y && (p.x == x || p.y == x);
    cnt[k] = mine ? a.n_pairs[ids[k]] : 0;
    side[k] = p.x == x ? 0 : 1;
    base[k] = mine ? a.edge_base[ids[k]] : 0;
  }
  const long long off = a.bs_off[x], first_elem_dof = a.elem_base[x];
  const long long n_elem_dofs = a.elem_base[a.n_elems];
  const int len = n_elem + cnt[0] + cnt[1] + cnt[2] + cnt[3];
  const int nu = ni * (nj - 1);
  for (int t = lane; t < len; t += kWave) {
    const long long pos = off + t;
    int dir, i, j, sd;
    long long dof;
    if (t < n_elem) {
      if (t < nu) {
        dir = 0;
        i = t / (nj - 1);
        j = 2 + t % (nj - 1);
      } else {
        const int u = t - nu;
        dir = 1;
        i = 2 + u / nj;
        j = u % nj;
      }
      dof = first_elem_dof + t;
      sd = 0;
    } else {
      int u = t - n_elem, k = 0;
      if (u >= cnt[0]) {
        u -= cnt[0];
        k = 1;
        if (u >= cnt[1]) {
          u -= cnt[1];
          k = 2;
          if (u >= cnt[2]) {
            u -= cnt[2];
            k = 3;
          }
        }
      }
      const int l = pick4(k, l0, l1, l2, l3);
      if (l < 2) {
        dir = 0;
        i = u;
        j = l;
      } else {
        dir = 1;
        i = l - 2;
        j = u;
      }
      dof = n_elem_dofs + pick4(k, base[0], base[1], base[2], base[3]) + u;
      sd = pick4(k, side[0], side[1], side[2], side[3]);
    }
    if (pos >= a.spec_cap || dof >= a.spec_cap) {
      atomicOr(&a.status->flags, kFlagCapacity);
      continue;
    }
    a.bs_dir[pos] = (unsigned char)dir;
    a.bs_i[pos] = (unsigned char)i;
    a.bs_j[pos] = (unsigned char)j;
    a.bs_dof[pos] = int(dof);
    a.dof_bs[2 * dof + sd] = int(pos);
    if (t < n_elem) a.dof_bs[2 * dof + 1] = -1;
  }
}

// Row groups: DoF d starts one when d = 0 or its (first Elem, second Elem)
// differs from DoF d - 1's
__device__ __forceinline__ int2 dof_elems(const int* dof_bs, const int* bs_elem, long long n_bs, long long d) {
  const int s0 = dof_bs[2 * d], s1 = dof_bs[2 * d + 1];
  return make_int2(s0 >= 0 && s0 < n_bs ? bs_elem[s0] : -1, s1 >= 0 && s1 < n_bs ? bs_elem[s1] : -1);
}

// grid-stride over the DoFs; one atomic per block
__global__ void __launch_bounds__(kBlock) k_group_count(const int* dof_bs, const int* bs_elem, long long n_dofs,
                                                        long long n_bs, Status* status) {
  __shared__ int wave_sum[kWavesPerBlock];
  int count = 0;
  for (long long d = (long long)blockIdx.x * kBlock + threadIdx.x; d < n_dofs; d += (long long)gridDim.x * kBlock) {
    if (d == 0) {
      ++count;
    } else {
      const int2 c = dof_elems(dof_bs, bs_elem, n_bs, d), p = dof_elems(dof_bs, bs_elem, n_bs, d - 1);
      count += c.x != p.x || c.y != p.y ? 1 : 0;
    }
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) count += __shfl_xor(count, off, kWave);
  if (int(threadIdx.x) % kWave == 0) wave_sum[threadIdx.x / kWave] = count;
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; ++i) total += wave_sum[i];
    if (total) atomicAdd(&status->groups, (unsigned long long)total);
  }
}

unsigned grid(long long n, long long per_block) { return unsigned((n + per_block - 1) / per_block); }

}  // namespace

long long scan_tiles(int n_elems, int n_edges) {
  auto t = [](long long n) { return n > 0 ? (n + kTile - 1) / kTile : 1; };
  return 2 * t(n_elems) + t(n_edges);
}

hipError_t launch_counts(const Args& a, hipStream_t st) {
  if (a.n_edges > 0) k_edges<<<grid(a.n_edges, kBlock), kBlock, 0, st>>>(a);
  if (a.n_elems > 0) k_elems<<<grid(a.n_elems, kBlock), kBlock, 0, st>>>(a);
  const unsigned tiles = unsigned(scan_tiles(a.n_elems, a.n_edges));
  k_scan_reduce<<<tiles, kBlock, 0, st>>>(a);
  k_scan_top<<<1, kBlock, 0, st>>>(a);
  k_scan_apply<<<tiles, kBlock, 0, st>>>(a);
  return hipGetLastError();
}

hipError_t launch_lists(const Args& a, hipStream_t st) {
  if (a.n_elems > 0) k_lists<<<grid(a.n_elems, kWavesPerBlock), kBlock, 0, st>>>(a);
  return hipGetLastError();
}

hipError_t launch_group_count(const int* dof_bs, const int* bs_elem, long long n_dofs, long long n_bs, Status* status,
                              hipStream_t st) {
  if (n_dofs > 0)
    k_group_count<<<unsigned(std::min<long long>(grid(n_dofs, kBlock), kGroupCountBlocks)), kBlock, 0, st>>>(
        dof_bs, bs_elem, n_dofs, n_bs, status);
  return hipGetLastError();
}

}  // namespace build
}  // namespace fem2d
