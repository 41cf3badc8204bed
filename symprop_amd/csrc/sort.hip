// Device-wide primitives for tree construction: exclusive scan and a stable
// LSD radix sort of packed sub-multiset keys (u64 or u128) with an optional
// u32 payload.  Wave64 throughout: in-wave ranking uses 64-bit ballots.
#include "sp_internal.hpp"

namespace spamd {

namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ inline u64 wave_incl_scan(u64 v) {
  const int lane = __lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    u64 t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Exclusive scan of the 256 per-thread totals of a block; returns this
// thread's exclusive prefix and writes the block total to *total.
__device__ inline u64 block_excl_scan(u64 v, u64 *total) {
  __shared__ u64 wsum[SCAN_BLOCK / 64];
  const int lane = __lane_id();
  const int w = threadIdx.x / 64;
  u64 inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  u64 off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < SCAN_BLOCK / 64; ++i) {
    if (i < w) off += wsum[i];
    tot += wsum[i];
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

__global__ void __launch_bounds__(SCAN_BLOCK)
scan_reduce_kernel(const u64 *__restrict__ in, u64 n, u64 *__restrict__ sums) {
  const u64 base = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_ITEMS;
  u64 s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i)
    if (base + i < n) s += in[base + i];
  u64 tot;
  (void)block_excl_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_BLOCK)
scan_block_kernel(const u64 *__restrict__ in, u64 *__restrict__ out, u64 n,
                  const u64 *__restrict__ block_off) {
  const u64 base = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_ITEMS;
  u64 v[SCAN_ITEMS];
  u64 s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0;
    s += v[i];
  }
  u64 tot;
  u64 run = block_excl_scan(s, &tot);
  if (block_off) run += block_off[blockIdx.x];
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
}

// ---------------------------------------------------------------- radix --
constexpr int RS_BLOCK = 256;
constexpr int RS_WAVES = RS_BLOCK / 64;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;

template <class K>
__device__ inline u32 digit_of(K k, int shift) {
  return (u32)(k >> shift) & 255u;
}

template <class K>
__global__ void __launch_bounds__(RS_BLOCK)
radix_hist_kernel(const K *__restrict__ keys, u64 n, int shift, u32 nblocks,
                  u64 *__restrict__ hist) {
  __shared__ u32 cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const u64 base = (u64)blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int i = 0; i < RS_ITEMS; ++i) {
    const u64 idx = base + (u64)i * RS_BLOCK + threadIdx.x;
    if (idx < n) atomicAdd(&cnt[digit_of(keys[idx], shift)], 1u);
  }
  __syncthreads();
  hist[(u64)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter: wave w owns the contiguous sub-tile
// [base + w*64*ITEMS, base + (w+1)*64*ITEMS), processed in rounds of 64
// consecutive keys, so (block, wave, round, lane) order equals input order.
template <class K>
__global__ void __launch_bounds__(RS_BLOCK)
radix_scatter_kernel(const K *__restrict__ kin, const u32 *__restrict__ vin,
                     K *__restrict__ kout, u32 *__restrict__ vout, u64 n,
                     int shift, u32 nblocks, const u64 *__restrict__ offs) {
  __shared__ u32 wcount[RS_WAVES][256];
  __shared__ u64 boff[256];
  const int lane = __lane_id();
  const int w = threadIdx.x / 64;
  for (int d = threadIdx.x; d < 256 * RS_WAVES; d += RS_BLOCK)
    (&wcount[0][0])[d] = 0;
  boff[threadIdx.x] = offs[(u64)threadIdx.x * nblocks + blockIdx.x];
  __syncthreads();

  const u64 base = (u64)blockIdx.x * RS_TILE + (u64)w * 64 * RS_ITEMS;
  const u64 ltmask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  K key[RS_ITEMS];
  u32 rank[RS_ITEMS];
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const u64 idx = base + (u64)r * 64 + lane;
    const bool valid = idx < n;
    key[r] = valid ? kin[idx] : (K)0;
    const u32 d = digit_of(key[r], shift);
    u64 peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const u64 bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    u32 prior = 0;
    if (valid) prior = wcount[w][d];
    const int leader = __ffsll((long long)peers) - 1;
    if (valid && lane == leader) wcount[w][d] = prior + (u32)__popcll(peers);
    rank[r] = prior + (u32)__popcll(peers & ltmask);
  }
  __syncthreads();
  {
    const int d = threadIdx.x;
    u32 run = 0;
#pragma unroll
    for (int i = 0; i < RS_WAVES; ++i) {
      const u32 t = wcount[i][d];
      wcount[i][d] = run;
      run += t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const u64 idx = base + (u64)r * 64 + lane;
    if (idx < n) {
      const u32 d = digit_of(key[r], shift);
      const u64 pos = boff[d] + wcount[w][d] + rank[r];
      kout[pos] = key[r];
      if (vin) vout[pos] = vin[idx];
    }
  }
}

}  // namespace

void exclusive_scan_u64(const u64 *in, u64 *out, u64 n, hipStream_t s) {
  if (n == 0) return;
  const u64 nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 1) {
    scan_block_kernel<<<1, SCAN_BLOCK, 0, s>>>(in, out, n, nullptr);
    SP_CHECK_LAUNCH();
    return;
  }
  DevBuf sums(nb * sizeof(u64));
  scan_reduce_kernel<<<(unsigned)nb, SCAN_BLOCK, 0, s>>>(in, n, sums.as<u64>());
  SP_CHECK_LAUNCH();
  exclusive_scan_u64(sums.as<u64>(), sums.as<u64>(), nb, s);
  scan_block_kernel<<<(unsigned)nb, SCAN_BLOCK, 0, s>>>(in, out, n, sums.as<u64>());
  SP_CHECK_LAUNCH();
  // sums is freed at scope exit; make sure the kernels that read it are done.
  SP_HIP(hipStreamSynchronize(s));
}

template <class K>
void radix_sort(K *keys, K *keys_tmp, u32 *vals, u32 *vals_tmp, u64 n,
                int bits, hipStream_t s, K **keys_out, u32 **vals_out) {
  K *ka = keys, *kb = keys_tmp;
  u32 *va = vals, *vb = vals_tmp;
  if (n > 1) {
    const u64 nb64 = (n + RS_TILE - 1) / RS_TILE;
    if (nb64 > 0xFFFFFFFFull) throw Error(SP_ERR_UNSUPPORTED, "sort too large");
    const u32 nb = (u32)nb64;
    DevBuf hist((u64)256 * nb * sizeof(u64));
    for (int shift = 0; shift < bits; shift += 8) {
      radix_hist_kernel<K><<<nb, RS_BLOCK, 0, s>>>(ka, n, shift, nb, hist.as<u64>());
      SP_CHECK_LAUNCH();
      exclusive_scan_u64(hist.as<u64>(), hist.as<u64>(), (u64)256 * nb, s);
      radix_scatter_kernel<K><<<nb, RS_BLOCK, 0, s>>>(ka, va, kb, vb, n, shift, nb,
                                                      hist.as<u64>());
      SP_CHECK_LAUNCH();
      std::swap(ka, kb);
      std::swap(va, vb);
    }
    SP_HIP(hipStreamSynchronize(s));
  }
  *keys_out = ka;
  if (vals_out) *vals_out = va;
}

template void radix_sort<u64>(u64 *, u64 *, u32 *, u32 *, u64, int, hipStream_t,
                              u64 **, u32 **);
template void radix_sort<u128>(u128 *, u128 *, u32 *, u32 *, u64, int,
                               hipStream_t, u128 **, u32 **);

}  // namespace spamd
