// Device-wide primitives of the map stage: a stable compaction and a stable
// LSD radix sort of (64-bit key, 32-bit value) pairs on a key's low bits.
// Both run as fixed launch sequences whose sizes come from the host and
// whose data-dependent counts stay on the device, and neither records
// anything but kernels, so the map stage can be captured as a hipGraph.
// Tiles of kPrimTile items.  The compaction gives thread t of a tile items
// t*kPrimIpt .. t*kPrimIpt + kPrimIpt - 1 (blocked order); the radix passes
// give wave w, lane l items w*64*kPrimIpt + k*64 + l for k = 0..kPrimIpt-1
// (wave-striped order: coalesced loads, and the wave-matching rank walks
// the items in index order), so in-tile ranks are stable either way.
#pragma once
#include <rocprim/block/block_radix_rank.hpp>
#include <rocprim/block/block_reduce.hpp>
#include <rocprim/block/block_scan.hpp>

namespace vmp {

constexpr int kPrimThreads = 256;
constexpr int kPrimIpt = 8;
constexpr int kPrimTile = kPrimThreads * kPrimIpt;
constexpr int kPrimScanThreads = 1024;
__host__ __device__ inline int prim_tiles(int64_t n) { return int((n + kPrimTile - 1) / kPrimTile); }
// int32 work entries the primitives need for n items: per-tile digit
// counts of a radix pass (256 per tile) plus two counters
__host__ __device__ inline size_t prim_ws_entries(int64_t n) { return size_t(prim_tiles(n)) * 256 + 64; }

// Exclusive scan in place of v[0..m) by one workgroup, in chunks of
// kPrimScanThreads * 8 entries: each chunk is loaded and stored coalesced
// through LDS and scanned in blocked order; the total goes to *total when
// total is not null.
constexpr int kPrimScanIpt = 8;
constexpr int kPrimScanChunk = kPrimScanThreads * kPrimScanIpt;
__global__ void __launch_bounds__(kPrimScanThreads) k_prim_scan(int32_t* v, int m, int32_t* total) {
  using Scan = rocprim::block_scan<int32_t, kPrimScanThreads>;
  __shared__ typename Scan::storage_type st;
  __shared__ int32_t buf[kPrimScanChunk];
  __shared__ int32_t carry_s;
  const int t = threadIdx.x;
  int32_t carry = 0;
  for (int c = 0; c < m; c += kPrimScanChunk) {
#pragma unroll
    for (int k = 0; k < kPrimScanIpt; ++k) {
      const int i = c + t + k * kPrimScanThreads;
      buf[t + k * kPrimScanThreads] = i < m ? v[i] : 0;
    }
    __syncthreads();
    int32_t x[kPrimScanIpt];
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kPrimScanIpt; ++k) {
      x[k] = buf[t * kPrimScanIpt + k];
      sum += x[k];
    }
    int32_t base = 0;
    Scan().exclusive_scan(sum, base, 0, st);
    base += carry;
#pragma unroll
    for (int k = 0; k < kPrimScanIpt; ++k) {
      buf[t * kPrimScanIpt + k] = base;
      base += x[k];
    }
    if (t == kPrimScanThreads - 1) carry_s = base;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPrimScanIpt; ++k) {
      const int i = c + t + k * kPrimScanThreads;
      if (i < m) v[i] = buf[t + k * kPrimScanThreads];
    }
    carry = carry_s;
    __syncthreads();
  }
  if (total && t == 0) *total = carry;
}

// ---- stable compaction: out[k] = val(i) for the k-th i < n with pred(i)
template <class Pred>
__global__ void __launch_bounds__(kPrimThreads) k_sel_count(int n, Pred pred, int32_t* tile_cnt) {
  using Reduce = rocprim::block_reduce<int32_t, kPrimThreads>;
  __shared__ typename Reduce::storage_type st;
  const int base = blockIdx.x * kPrimTile + threadIdx.x * kPrimIpt;
  int32_t c = 0;
#pragma unroll
  for (int k = 0; k < kPrimIpt; ++k) c += (base + k < n && pred(base + k)) ? 1 : 0;
  int32_t tot;
  Reduce().reduce(c, tot, st);
  if (threadIdx.x == 0) tile_cnt[blockIdx.x] = tot;
}
template <class Pred, class Out>
__global__ void __launch_bounds__(kPrimThreads) k_sel_scatter(int n, Pred pred, const int32_t* tile_off, Out out) {
  using Scan = rocprim::block_scan<int32_t, kPrimThreads>;
  __shared__ typename Scan::storage_type st;
  const int base = blockIdx.x * kPrimTile + threadIdx.x * kPrimIpt;
  bool f[kPrimIpt];
  int32_t c = 0;
#pragma unroll
  for (int k = 0; k < kPrimIpt; ++k) {
    f[k] = base + k < n && pred(base + k);
    c += f[k] ? 1 : 0;
  }
  int32_t pos = 0;
  Scan().exclusive_scan(c, pos, 0, st);
  pos += tile_off[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kPrimIpt; ++k)
    if (f[k]) out(pos++, base + k);
}

// flags[i] != 0 -> the ascending indices, as rocprim::select over a counting
// iterator would give them
struct PredFlag {
  const int32_t* f;
  __device__ bool operator()(int i) const { return f[i] != 0; }
};
struct OutIndex {
  int32_t* o;
  __device__ void operator()(int k, int i) const { o[k] = i; }
};

// stable compaction; work: prim_ws_entries(n) int32 entries, *count the total
template <class Pred, class Out>
inline void select_stable(hipStream_t st, int n, Pred pred, Out out, int32_t* work, int32_t* count) {
  const int nt = prim_tiles(n);
  if (nt == 0) {
    hipLaunchKernelGGL(k_prim_scan, dim3(1), dim3(kPrimScanThreads), 0, st, work, 0, count);
    return;
  }
  hipLaunchKernelGGL((k_sel_count<Pred>), dim3(nt), dim3(kPrimThreads), 0, st, n, pred, work);
  hipLaunchKernelGGL(k_prim_scan, dim3(1), dim3(kPrimScanThreads), 0, st, work, nt, count);
  hipLaunchKernelGGL((k_sel_scatter<Pred, Out>), dim3(nt), dim3(kPrimThreads), 0, st, n, pred, work, out);
}

// largest input the primitives take (the tile scans run in one workgroup);
// larger batches use rocPRIM's device-wide routines
constexpr int kPrimMaxItems = 1 << 20;

// ---- stable LSD radix sort of (key, value) on key bits [0, bits)
// digit histogram of each tile, stored digit-major: hist[d * ntile + tile]
template <class K>
__global__ void __launch_bounds__(kPrimThreads) k_rs_hist(const K* keys, int n, int shift, uint32_t mask,
                                                         int32_t* hist) {
  __shared__ int32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * kPrimTile + (threadIdx.x >> 6) * 64 * kPrimIpt + (threadIdx.x & 63);
#pragma unroll
  for (int k = 0; k < kPrimIpt; ++k)
    if (base + k * 64 < n) atomicAdd(&cnt[uint32_t(keys[base + k * 64] >> shift) & mask], 1);
  __syncthreads();
  hist[size_t(threadIdx.x) * gridDim.x + blockIdx.x] = cnt[threadIdx.x];
}
// stable scatter of one pass: in-tile ranks by digit (blocked order), the
// tile sorted by digit in LDS, then written in LDS order so that each
// digit's run goes out contiguously (coalesced) at that digit's offset for
// this tile from the scanned histogram
template <class K>
__global__ void __launch_bounds__(kPrimThreads) k_rs_scatter(const K* kin, const int32_t* vin, K* kout, int32_t* vout,
                                                            int n, int shift, uint32_t mask, const int32_t* offs) {
  using Rank = rocprim::block_radix_rank<kPrimThreads, 8, rocprim::block_radix_rank_algorithm::match>;
  static_assert(Rank::digits_per_thread == 1, "one digit per thread");
  __shared__ typename Rank::storage_type st;
  __shared__ int32_t dstart[256];  // first in-tile rank of each digit
  __shared__ int32_t dbase[256];   // the digit's output offset for this tile
  __shared__ K skey[kPrimTile];
  __shared__ int32_t sval[kPrimTile];
  const int tile0 = blockIdx.x * kPrimTile;
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
  dstart[threadIdx.x] = int32_t(prefix[0]);
  dbase[threadIdx.x] = offs[size_t(threadIdx.x) * gridDim.x + blockIdx.x];
#pragma unroll
  for (int k = 0; k < kPrimIpt; ++k) {
    skey[rank[k]] = key[k];
    sval[rank[k]] = val[k];
  }
  __syncthreads();
  // the valid items occupy ranks 0..valid-1 (invalid ones rank last)
  for (int i = threadIdx.x; i < valid; i += kPrimThreads) {
    const K kk = skey[i];
    const uint32_t d = uint32_t(kk >> shift) & mask;
    const int pos = dbase[d] + (i - dstart[d]);
    kout[pos] = kk;
    vout[pos] = sval[i];
  }
}

// passes for `bits` key bits: an odd number (so the result lands in the
// output buffers) of at most 8 bits each
inline int rs_passes(int bits) {
  int p = bits <= 0 ? 1 : (bits + 7) / 8;
  if ((p & 1) == 0) ++p;
  return p;
}
// Sorts (kin, vin) into (kout, vout) by the low `bits` bits of the keys,
// stably; kin / vin are overwritten (ping-pong buffers).  work:
// prim_ws_entries(n) int32 entries.
template <class K>
inline void radix_sort_stable(hipStream_t st, K* kin, int32_t* vin, K* kout, int32_t* vout, int n, int bits,
                              int32_t* work) {
  const int nt = prim_tiles(n);
  if (nt == 0) return;
  const int passes = rs_passes(bits);
  const int width = bits <= 0 ? 1 : (bits + passes - 1) / passes;
  const uint32_t mask = (1u << width) - 1u;
  for (int p = 0; p < passes; ++p) {
    const bool fwd = (p & 1) == 0;  // even passes in -> out, odd ones out -> in
    const K* ks = fwd ? kin : kout;
    const int32_t* vs = fwd ? vin : vout;
    K* kd = fwd ? kout : kin;
    int32_t* vd = fwd ? vout : vin;
    const int shift = p * width;
    hipLaunchKernelGGL(k_rs_hist<K>, dim3(nt), dim3(kPrimThreads), 0, st, ks, n, shift, mask, work);
    hipLaunchKernelGGL(k_prim_scan, dim3(1), dim3(kPrimScanThreads), 0, st, work, nt * 256, nullptr);
    hipLaunchKernelGGL(k_rs_scatter<K>, dim3(nt), dim3(kPrimThreads), 0, st, ks, vs, kd, vd, n, shift, mask, work);
  }
}

}  // namespace vmp
