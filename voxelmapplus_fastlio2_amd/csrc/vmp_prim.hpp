// Device-wide primitives of the map stage: a stable compaction (one launch)
// and a stable LSD radix sort of (key, 32-bit value) pairs on a key's low
// bits (one histogram launch, then one launch per pass).  Their sizes come
// from the host, their data-dependent counts stay on the device, and they
// launch nothing but kernels, so the map stage can be captured as a hipGraph.
// Tiles of kPrimTile items.  The compaction gives thread t of a tile items
// t*kPrimIpt .. t*kPrimIpt + kPrimIpt - 1 (blocked order); the radix passes
// give wave w, lane l items w*64*kPrimIpt + k*64 + l for k = 0..kPrimIpt-1
// (wave-striped order: coalesced loads, and the wave-matching rank walks
// the items in index order), so in-tile ranks are stable either way.
#pragma once
#include <algorithm>

#include <rocprim/block/block_radix_rank.hpp>
#include <rocprim/block/block_reduce.hpp>
#include <rocprim/block/block_scan.hpp>

namespace vmp {

constexpr int kPrimThreads = 256;
constexpr int kPrimIpt = 8;
constexpr int kPrimTile = kPrimThreads * kPrimIpt;
__host__ __device__ inline int prim_tiles(int64_t n) { return int((n + kPrimTile - 1) / kPrimTile); }
// ---- single-pass control.  A launch takes its tiles in ticket order (a
// tile's predecessors have all started, so waiting on them always ends),
// publishes per tile its aggregate and then its inclusive prefix in status
// words (bits 30-31: 1 aggregate, 2 inclusive; bits 0-29 the value), and
// finds its own prefix by looking back over the predecessors' words.  The
// tile that takes the last ticket resets the ticket counter; the last tile to
// finish clears the status words and the done counter, so every launch
// starts from zeros (the work area is zeroed once when it is allocated).
struct PrimCtl {
  int32_t ticket;
  int32_t done;
  int32_t err;  // a bounded wait ran out (not expected; the results are then void)
  int32_t pad[61];
};
constexpr uint32_t kStAgg = 1u << 30, kStInc = 2u << 30, kStVal = (1u << 30) - 1;
constexpr int kSpinMax = 1 << 22;
constexpr int kPrimMaxPasses = 9;  // radix passes of up to 64 key bits (odd count)
// work area of n items, in int32 entries: control, the radix histograms of
// every pass, then 256 status words per tile
__host__ __device__ inline size_t prim_ws_entries(int64_t n) {
  return 64 + size_t(kPrimMaxPasses) * 256 + size_t(prim_tiles(n)) * 256;
}
struct PrimWs {
  PrimCtl* ctl;
  int32_t* ghist;    // [kPrimMaxPasses][256]
  uint32_t* status;  // [tiles][256] (the compaction uses [tiles])
};
__host__ __device__ inline PrimWs prim_ws(int32_t* w) {
  return PrimWs{reinterpret_cast<PrimCtl*>(w), w + 64, reinterpret_cast<uint32_t*>(w + 64 + kPrimMaxPasses * 256)};
}

__device__ inline uint32_t st_load(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_store(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// the tile's ticket (thread 0; broadcast by the caller)
__device__ inline int take_ticket(PrimCtl* ctl, int nt) {
  const int t = atomicAdd(&ctl->ticket, 1);
  if (t == nt - 1) atomicExch(&ctl->ticket, 0);
  return t;
}
// Look-back: the exclusive prefix of a tile from its predecessors' status
// words, read a window at a time (the words of one window are loaded
// together, so a chain of aggregates costs one memory round trip per window
// rather than one per tile).  The ready words nearest the tile are consumed
// up to the first inclusive one; a window whose nearest word is not yet
// published is read again.  A bounded wait that runs out sets ctl->err.
//
// One thread per slot, a window of kLbWin predecessors (the radix passes:
// slot = digit, stride 256)
constexpr int kLbWin = 8;
__device__ inline uint32_t look_back(const uint32_t* status, int tile, int stride, int slot, PrimCtl* ctl) {
  uint32_t prefix = 0;
  int j = tile - 1, spins = 0;
  while (j >= 0) {
    uint32_t w[kLbWin];
#pragma unroll
    for (int k = 0; k < kLbWin; ++k) w[k] = j - k >= 0 ? st_load(status + size_t(j - k) * stride + slot) : kStInc;
    bool stop = false, done = false;
    int adv = 0;
#pragma unroll
    for (int k = 0; k < kLbWin; ++k) {
      if (!stop) {
        if ((w[k] >> 30) == 0) {
          stop = true;
        } else {
          prefix += w[k] & kStVal;
          ++adv;
          if ((w[k] & kStInc) == kStInc) stop = done = true;
        }
      }
    }
    if (done) break;
    if (adv == 0) {
      if (++spins > kSpinMax) {
        atomicOr(&ctl->err, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    j -= adv;
  }
  return prefix;
}
__device__ inline uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += uint32_t(__shfl_xor(int(v), o, 64));
  return v;
}
// One wave, lane l reading predecessor j - l (the compaction: one word per
// tile); every lane of the wave calls it and gets the prefix.
__device__ inline uint32_t look_back_wave(const uint32_t* status, int tile, PrimCtl* ctl) {
  const int lane = threadIdx.x & 63;
  uint32_t prefix = 0;
  int j = tile - 1, spins = 0;
  while (j >= 0) {
    const int idx = j - lane;
    const uint32_t w = idx >= 0 ? st_load(status + idx) : kStInc;  // before tile 0: an inclusive 0
    const uint64_t nready = __ballot((w >> 30) == 0);
    const uint64_t inc = __ballot((w & kStInc) == kStInc);
    const int nr = nready ? __ffsll((unsigned long long)nready) - 1 : 64;
    const int fi = inc ? __ffsll((unsigned long long)inc) - 1 : 64;
    if (fi < nr) {
      prefix += wave_sum_u32(lane <= fi ? (w & kStVal) : 0u);
      break;
    }
    if (nr == 0) {
      if (++spins > kSpinMax) {
        if (lane == 0) atomicOr(&ctl->err, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    prefix += wave_sum_u32(lane < nr ? (w & kStVal) : 0u);
    j -= nr;
  }
  return prefix;
}
// end of a tile: the last to finish clears `nclear` status words (and the
// extra words, if any) and the done counter; returns true in that tile
// (every thread of it).  With `publish`, the tile's stores are released
// device-wide before it counts itself done (the workgroup barrier orders
// every thread's stores before thread 0's agent-scope release fence, which
// the memory model makes cumulative), and the last tile acquires them, so it
// may read what every tile wrote.
__device__ inline bool finish_tile(PrimCtl* ctl, int nt, uint32_t* status, int nclear, int32_t* extra, int nextra,
                                   int* s_last, bool publish = false) {
  __syncthreads();
  if (threadIdx.x == 0) {
    if (publish) __threadfence();
    *s_last = atomicAdd(&ctl->done, 1) == nt - 1;
    if (publish && *s_last) __threadfence();
  }
  __syncthreads();
  if (!*s_last) return false;
  for (int i = threadIdx.x; i < nclear; i += blockDim.x) status[i] = 0u;
  for (int i = threadIdx.x; i < nextra; i += blockDim.x) extra[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) atomicExch(&ctl->done, 0);
  return true;
}

// Hooks of a compaction: at_tile(tile, base, pos, v) runs in every thread
// once the tile's prefix is known (its items base .. base + kPrimIpt - 1
// with values v, pos the output position of its first selected item);
// operator()(count) runs in the last tile to finish (kActive).
struct NoTail {
  static constexpr bool kActive = false;
  __device__ void at_tile(int, int, int, const int32_t (&)[kPrimIpt]) const {}
  __device__ void operator()(int) const {}
};

// ---- stable compaction: for the k-th i < n (ascending) whose pred(i) = v is
// non-zero, out(k, i, v).  pred is called once per item and may clear what it
// read.  tail(count), if active, runs in every thread of the last tile to
// finish, after the outputs of all tiles are visible.
template <class Pred, class Out, class Tail>
__global__ void __launch_bounds__(kPrimThreads) k_select1(int n, Pred pred, Out out, Tail tail, int32_t* count,
                                                          PrimWs ws) {
  using Scan = rocprim::block_scan<int32_t, kPrimThreads>;
  __shared__ typename Scan::storage_type st;
  __shared__ int s_tile, s_last;
  __shared__ uint32_t s_prefix;
  const int nt = gridDim.x;
  if (threadIdx.x == 0) s_tile = take_ticket(ws.ctl, nt);
  __syncthreads();
  const int tile = s_tile;
  const int base = tile * kPrimTile + threadIdx.x * kPrimIpt;
  int32_t v[kPrimIpt];
  int32_t c = 0;
#pragma unroll
  for (int k = 0; k < kPrimIpt; ++k) {
    v[k] = base + k < n ? pred(base + k) : 0;
    c += v[k] != 0 ? 1 : 0;
  }
  int32_t pos = 0, agg = 0;
  Scan().exclusive_scan(c, pos, 0, agg, st);
  if (threadIdx.x < 64) {
    const bool l0 = threadIdx.x == 0;
    uint32_t prefix = 0;
    if (tile == 0) {
      if (l0) st_store(ws.status, kStInc | uint32_t(agg));
    } else {
      if (l0) st_store(ws.status + tile, kStAgg | uint32_t(agg));
      prefix = look_back_wave(ws.status, tile, ws.ctl);
      if (l0) st_store(ws.status + tile, kStInc | (prefix + uint32_t(agg)));
    }
    if (l0) s_prefix = prefix;
  }
  __syncthreads();
  pos += int32_t(s_prefix);
  tail.at_tile(tile, base, pos, v);
#pragma unroll
  for (int k = 0; k < kPrimIpt; ++k)
    if (v[k] != 0) out(pos++, base + k, v[k]);
  if (threadIdx.x == 0 && tile == nt - 1) *count = int32_t(s_prefix) + agg;
  if (finish_tile(ws.ctl, nt, ws.status, nt, nullptr, 0, &s_last, Tail::kActive) && Tail::kActive) tail(*count);
}

// flags[i] != 0 -> the ascending indices, as rocprim::select over a counting
// iterator would give them
struct PredFlag {
  const int32_t* f;
  __device__ int32_t operator()(int i) const { return f[i]; }
};
// the same, leaving the flags it read cleared
struct PredFlagClear {
  int32_t* f;
  __device__ int32_t operator()(int i) const {
    const int32_t v = f[i];
    if (v != 0) f[i] = 0;
    return v;
  }
};
struct OutIndex {
  int32_t* o;
  __device__ void operator()(int k, int i, int32_t) const { o[k] = i; }
};

// stable compaction, one launch; work: prim_ws_entries(n) int32 entries,
// *count the total
template <class Pred, class Out, class Tail = NoTail>
inline void select_stable(hipStream_t st, int n, Pred pred, Out out, int32_t* work, int32_t* count,
                          Tail tail = Tail()) {
  const int nt = std::max(1, prim_tiles(n));
  hipLaunchKernelGGL((k_select1<Pred, Out, Tail>), dim3(nt), dim3(kPrimThreads), 0, st, n, pred, out, tail, count,
                     prim_ws(work));
}

// largest input the primitives take; larger batches use rocPRIM's
// device-wide routines
constexpr int kPrimMaxItems = 1 << 20;

// ---- stable LSD radix sort of (key, value) on key bits [0, bits)
// the digit histograms of every pass over the whole input, added into
// ghist[pass][256] (which the sort's last pass clears again)
template <class K>
__global__ void __launch_bounds__(kPrimThreads) k_rs_ghist(const K* keys, int n, int passes, int width,
                                                          int32_t* ghist) {
  __shared__ int32_t h[kPrimMaxPasses * 256];
  for (int i = threadIdx.x; i < passes * 256; i += kPrimThreads) h[i] = 0;
  __syncthreads();
  const uint32_t mask = (1u << width) - 1u;
  const int base = blockIdx.x * kPrimTile + (threadIdx.x >> 6) * 64 * kPrimIpt + (threadIdx.x & 63);
#pragma unroll
  for (int k = 0; k < kPrimIpt; ++k) {
    if (base + k * 64 >= n) continue;
    const K key = keys[base + k * 64];
    for (int p = 0; p < passes; ++p) atomicAdd(&h[p * 256 + (uint32_t(key >> (p * width)) & mask)], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * 256; i += kPrimThreads)
    if (h[i]) atomicAdd(&ghist[i], h[i]);
}
// One pass: in-tile ranks by digit (wave matching over wave-striped items),
// each digit's offset = its global start (scan of the pass's histogram) +
// the digit's count in earlier tiles (look-back per digit, one thread per
// digit), the tile sorted by digit in LDS and written out coalesced.
template <class K>
__global__ void __launch_bounds__(kPrimThreads) k_rs_pass(const K* kin, const int32_t* vin, K* kout, int32_t* vout,
                                                         int n, int shift, uint32_t mask, const int32_t* hist,
                                                         PrimWs ws, int clear_hist) {
  using Rank = rocprim::block_radix_rank<kPrimThreads, 8, rocprim::block_radix_rank_algorithm::match>;
  using Scan = rocprim::block_scan<int32_t, kPrimThreads>;
  static_assert(Rank::digits_per_thread == 1, "one digit per thread");
  __shared__ typename Rank::storage_type st;
  __shared__ typename Scan::storage_type sst;
  __shared__ int32_t dstart[256];  // first in-tile rank of each digit
  __shared__ int32_t dbase[256];   // the digit's output offset for this tile
  __shared__ K skey[kPrimTile];
  __shared__ int32_t sval[kPrimTile];
  __shared__ int s_tile, s_last;
  const int nt = gridDim.x;
  if (threadIdx.x == 0) s_tile = take_ticket(ws.ctl, nt);
  __syncthreads();
  const int tile = s_tile;
  const int tile0 = tile * kPrimTile;
  const int base = tile0 + (threadIdx.x >> 6) * 64 * kPrimIpt + (threadIdx.x & 63);
  const int valid = min(kPrimTile, n - tile0);
  K key[kPrimIpt];
  int32_t val[kPrimIpt];
#pragma unroll
  for (int k = 0; k < kPrimIpt; ++k) {
    const bool in = base + k * 64 < n;
    // items past n carry the largest digit and the tile's largest indices,
    // so they rank after every valid item and are not written
    key[k] = in ? kin[base + k * 64] : ~K(0);
    val[k] = in ? vin[base + k * 64] : 0;
  }
  unsigned int rank[kPrimIpt];
  unsigned int prefix[1], counts[1];
  Rank().rank_keys(
      key, rank, st, [&](const K& x) { return uint32_t(x >> shift) & mask; }, prefix, counts);
  const int d = threadIdx.x;
  // this tile's count of digit d (the invalid items, all of digit `mask`,
  // are not counted)
  const uint32_t cnt_d = counts[0] - (d == int(mask) ? uint32_t(kPrimTile - valid) : 0u);
  uint32_t* sd = ws.status + d;
  uint32_t before = 0;
  if (tile == 0) {
    st_store(sd, kStInc | cnt_d);
  } else {
    st_store(sd + size_t(tile) * 256, kStAgg | cnt_d);
    before = look_back(ws.status, tile, 256, d, ws.ctl);
    st_store(sd + size_t(tile) * 256, kStInc | (before + cnt_d));
  }
  int32_t gstart = 0;
  Scan().exclusive_scan(hist[d], gstart, 0, sst);
  dstart[d] = int32_t(prefix[0]);
  dbase[d] = gstart + int32_t(before);
#pragma unroll
  for (int k = 0; k < kPrimIpt; ++k) {
    skey[rank[k]] = key[k];
    sval[rank[k]] = val[k];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < valid; i += kPrimThreads) {
    const K kk = skey[i];
    const uint32_t dd = uint32_t(kk >> shift) & mask;
    kout[dbase[dd] + (i - dstart[dd])] = kk;
    vout[dbase[dd] + (i - dstart[dd])] = sval[i];
  }
  finish_tile(ws.ctl, nt, ws.status, nt * 256, clear_hist ? ws.ghist : nullptr,
              clear_hist ? kPrimMaxPasses * 256 : 0, &s_last);
}

// passes for `bits` key bits: an odd number (so the result lands in the
// output buffers) of at most 8 bits each
inline int rs_passes(int bits) {
  int p = bits <= 0 ? 1 : (bits + 7) / 8;
  if ((p & 1) == 0) ++p;
  return p;
}
// Sorts (kin, vin) into (kout, vout) by the low `bits` bits of the keys,
// stably; kin / vin are overwritten (ping-pong buffers).  One histogram
// launch, then one launch per pass.  work: prim_ws_entries(n) int32 entries.
template <class K>
inline void radix_sort_stable(hipStream_t st, K* kin, int32_t* vin, K* kout, int32_t* vout, int n, int bits,
                              int32_t* work) {
  const int nt = prim_tiles(n);
  if (nt == 0) return;
  const int passes = rs_passes(bits);
  const int width = bits <= 0 ? 1 : (bits + passes - 1) / passes;
  const uint32_t mask = (1u << width) - 1u;
  const PrimWs ws = prim_ws(work);
  hipLaunchKernelGGL(k_rs_ghist<K>, dim3(nt), dim3(kPrimThreads), 0, st, kin, n, passes, width, ws.ghist);
  for (int p = 0; p < passes; ++p) {
    const bool fwd = (p & 1) == 0;  // even passes in -> out, odd ones out -> in
    const K* ks = fwd ? kin : kout;
    const int32_t* vs = fwd ? vin : vout;
    K* kd = fwd ? kout : kin;
    int32_t* vd = fwd ? vout : vin;
    hipLaunchKernelGGL(k_rs_pass<K>, dim3(nt), dim3(kPrimThreads), 0, st, ks, vs, kd, vd, n, p * width, mask,
                       ws.ghist + p * 256, ws, p == passes - 1 ? 1 : 0);
  }
}

}  // namespace vmp
