// Compressed dimension tree construction on the GPU.
//
// The reference builds a pointer trie by inserting, for every nonzero M
// (sorted, 1-based) and j = N-1..0, every sub-multiset of M \ {a_j} and an
// owner (rem = a_j, value) at the leaf M \ {a_j} (src/dtree.cpp:38-80), then
// flattens it breadth-first with children sorted by index
// (src/compressed_dtree.cpp:13-90) and computes exclude-one parents
// (src/compressed_dtree.cpp:92-131).  The flattened result is fixed by
// sets and orders alone:
//   level l      = the distinct sub-multisets of size l+1 (l+1 <= N-1) of the
//                  nonzeros, in lexicographic order;
//   level_idx    = last index of each node;
//   level_ptr[l] = CSR from level l to level l+1 (children = nodes whose
//                  first l+1 indices equal the parent) and, at the leaf level,
//                  into the owner list;
//   owners       = (a_j, v) at leaf M \ {a_j}, ordered by leaf, then
//                  nonzero input order, then j descending;
//   dep_list[l]  = level-(l-1) index of node j with position k removed.
// Here each level is produced by packing every size-(l+1) position subset
// of every nonzero into an integer key (b bits per index, first index most
// significant, so integer order = lexicographic order), a stable LSD radix
// sort, and unique/segment extraction.  Parents and dep_list come from
// binary searches over the sorted keys of the level below.
#include "sp_internal.hpp"

#include <algorithm>
#include <cstring>

namespace spamd {

namespace {

constexpr int BLK = 256;

template <class K>
__device__ inline K pack_tuple(const u32 *t, const int *pos, int s, int b) {
  K k = 0;
  for (int i = 0; i < s; ++i) k = (k << b) | (K)t[pos[i]];
  return k;
}

// Keys of all size-s position subsets of all nonzeros.  comb: ncomb x s.
template <class K>
__global__ void gen_level_keys(const u32 *__restrict__ inds, u64 nnz, int N,
                               const int *__restrict__ comb, int ncomb, int s,
                               int b, K *__restrict__ out) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnz * (u64)ncomb) return;
  const u64 i = t / ncomb;
  const int c = (int)(t % ncomb);
  out[t] = pack_tuple<K>(inds + i * N, comb + c * s, s, b);
}

// Owner records: record id i*N + o, o = 0..N-1 <-> j = N-1-o; key = leaf.
template <class K>
__global__ void gen_owner_keys(const u32 *__restrict__ inds, u64 nnz, int N,
                               int b, K *__restrict__ keys, u32 *__restrict__ rec) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnz * (u64)N) return;
  const u64 i = t / N;
  const int j = N - 1 - (int)(t % N);
  const u32 *m = inds + i * N;
  K k = 0;
  for (int q = 0; q < N; ++q)
    if (q != j) k = (k << b) | (K)m[q];
  keys[t] = k;
  rec[t] = (u32)t;
}

template <class K>
__global__ void flag_heads(const K *__restrict__ keys, u64 n, u64 *__restrict__ flag) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// pos = exclusive scan of head flags; writes unique keys, and the segment
// id of every element (if seg != null) and segment starts (if start != null).
template <class K>
__global__ void compact_heads(const K *__restrict__ keys, u64 n,
                              const u64 *__restrict__ pos, K *__restrict__ uniq,
                              u64 *__restrict__ seg, u64 *__restrict__ start) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool head = (i == 0 || keys[i] != keys[i - 1]);
  // pos[i] counts heads before i, so a head's own id is pos[i]
  const u64 id = head ? pos[i] : pos[i] - 1;
  if (head) {
    uniq[id] = keys[i];
    if (start) start[id] = i;
  }
  if (seg) seg[i] = id;
}

template <class K>
__global__ void last_index(const K *__restrict__ keys, u64 n, int b,
                           u32 *__restrict__ idx) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  idx[i] = (u32)(keys[i] & (K)((1ull << b) - 1));
}

template <class K>
__device__ inline u64 lower_bound_key(const K *__restrict__ a, u64 n, K x) {
  u64 lo = 0, hi = n;
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// CSR from parents to children: the children of parent key k are the child
// keys in [k << b, (k+1) << b), so ptr[i] = lower_bound(children, key_i << b)
// (children with smaller prefixes sort first).  ptr[np] = nc.
// The search runs within the children sharing the parent's leading field
// (cfidx: first_field_index of the child level; sp = parent fields).
template <class KC, class KP>
__global__ void child_ptr(const KC *__restrict__ child, u64 nc, int b,
                          const KP *__restrict__ parent, u64 np, int sp,
                          const u64 *__restrict__ cfidx, u64 *__restrict__ ptr) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > np) return;
  if (i == np) {
    ptr[i] = nc;
    return;
  }
  const KP key = parent[i];
  const u32 f = (u32)(key >> (b * (sp - 1)));
  const u64 lo = cfidx[f];
  ptr[i] = lo + lower_bound_key<KC>(child + lo, cfidx[f + 1] - lo, (KC)key << b);
}

// idx[v] = first key of the level whose leading field is >= v, v = 0..dim+1
// (a level's keys hold fields of b bits, indices 1..dim, leading field most
// significant); idx[dim+1] = n.
template <class K>
__global__ void first_field_index(const K *__restrict__ keys, u64 n, int fields, int b,
                                  u32 dim, u64 *__restrict__ idx) {
  const u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > (u64)dim + 1) return;
  idx[v] = v > dim ? n : lower_bound_key<K>(keys, n, (K)v << (b * (fields - 1)));
}

// dep_list: node j (s fields) minus position k, searched in the level below
// within the keys that share its leading field (fidx of the level below).
template <class KC, class KP>
__global__ void exclude_one(const KC *__restrict__ keys, u64 n, int s, int b,
                            const KP *__restrict__ below, const u64 *__restrict__ fidx,
                            u64 *__restrict__ dep) {
  const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * (u64)s) return;
  const u64 j = t / s;
  const int k = (int)(t % s);
  const KC key = keys[j];
  const int lo_bits = b * (s - 1 - k);
  const KC lo_mask = lo_bits ? (((KC)1 << lo_bits) - 1) : (KC)0;
  // fields 0..k-1; for k == 0 the shift would equal the key width
  const KC hi = (k == 0) ? (KC)0 : (key >> (b * (s - k)));
  const KC sub = (hi << lo_bits) | (key & lo_mask);
  const u32 f = (u32)(sub >> (b * (s - 2)));
  const u64 lo = fidx[f];
  dep[t] = lo + lower_bound_key<KP>(below + lo, fidx[f + 1] - lo, (KP)sub);
}

__global__ void owner_fields(const u32 *__restrict__ rec, u64 no, int N,
                             const u32 *__restrict__ inds,
                             const double *__restrict__ vals,
                             u32 *__restrict__ rem, double *__restrict__ val) {
  const u64 p = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= no) return;
  const u32 r = rec[p];
  const u64 i = r / N;
  const int j = N - 1 - (int)(r % N);
  rem[p] = inds[i * N + j];
  val[p] = vals[i];
}

__global__ void gather_ptr_tail(u64 *__restrict__ ptr, u64 L, u64 total) {
  ptr[L] = total;
}

// rem-sorted work list: tuple of the owner's leaf (0-based rows) and
// coefficient v * N!/prod f! / N! computed as in
// src/compressed_dtree.cpp:493-503.
template <class K>
__global__ void build_worklist(const u32 *__restrict__ order_rec,  // owner pos
                               u64 no, const u64 *__restrict__ own_leaf,
                               const u32 *__restrict__ own_rem,
                               const double *__restrict__ own_val,
                               const K *__restrict__ leaf_keys, int m, int b,
                               u32 *__restrict__ tup, double *__restrict__ coef,
                               float *__restrict__ coef32, double *__restrict__ val,
                               u32 *__restrict__ leaf) {
  const u64 q = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= no) return;
  const u32 o = order_rec[q];
  const u64 lid = own_leaf[o];
  leaf[q] = (u32)lid;
  const K key = leaf_keys[lid];
  const u32 rem = own_rem[o];
  u32 t[16];
  const K mask = (K)((1ull << b) - 1);
  for (int i = 0; i < m; ++i) t[i] = (u32)((key >> (b * (m - 1 - i))) & mask);
  for (int i = 0; i < m; ++i) tup[q * m + i] = t[i] - 1;
  // multinomial of the full sorted tuple leaf + {rem}
  u64 fact[20];
  fact[0] = 1;
  for (int i = 1; i < 20; ++i) fact[i] = fact[i - 1] * (u64)i;
  u64 denom = 1;
  int run = 1;
  // insert rem into the sorted leaf to get the sorted nonzero tuple
  u32 full[17];
  int w = 0;
  bool placed = false;
  for (int i = 0; i < m; ++i) {
    if (!placed && rem <= t[i]) { full[w++] = rem; placed = true; }
    full[w++] = t[i];
  }
  if (!placed) full[w++] = rem;
  for (int i = 1; i <= m + 1; ++i) {
    if (i <= m && full[i] == full[i - 1]) {
      ++run;
    } else {
      denom *= fact[run];
      run = 1;
    }
  }
  const u64 mult = fact[m + 1] / denom;
  const double c = own_val[o] * (double)mult / (double)fact[m + 1];
  coef[q] = c;
  coef32[q] = (float)c;
  val[q] = own_val[o];
}

__global__ void rem_keys(const u32 *__restrict__ rem, u64 no, u64 *__restrict__ key,
                         u32 *__restrict__ id) {
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= no) return;
  key[i] = rem[i];
  id[i] = (u32)i;
}

// Row segments of the rem-sorted owners: segptr[r] = first owner with
// rem >= r+1 (rem is 1-based), segptr[dim] = owners.
__global__ void seg_bounds(const u64 *__restrict__ sorted_rem, u64 no, u32 dim,
                           u64 *__restrict__ segptr) {
  const u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > dim) return;
  segptr[r] = (r == dim) ? no : lower_bound_key<u64>(sorted_rem, no, r + 1);
}

std::vector<int> combinations(int N, int s) {
  std::vector<int> out;
  std::vector<int> c(s);
  for (int i = 0; i < s; ++i) c[i] = i;
  while (true) {
    out.insert(out.end(), c.begin(), c.end());
    int p = s - 1;
    while (p >= 0 && c[p] == N - s + p) --p;
    if (p < 0) break;
    c[p]++;
    for (int q = p + 1; q < s; ++q) c[q] = c[q - 1] + 1;
  }
  return out;
}

int key_words_for(int s, int b) { return s * b <= 64 ? 1 : 2; }

// Sorted unique keys of one level (interior) built from all position subsets.
template <class K>
void build_interior_level(Ctx &c, int l) {
  const int N = (int)c.order, s = l + 1, b = c.bits;
  const std::vector<int> comb = combinations(N, s);
  const int ncomb = (int)(comb.size() / s);
  const u64 cnt = c.nnz * (u64)ncomb;
  DevBuf d_comb(comb.size() * sizeof(int));
  SP_HIP(hipMemcpyAsync(d_comb.p, comb.data(), comb.size() * sizeof(int),
                        hipMemcpyHostToDevice, c.stream));
  DevBuf ka(cnt * sizeof(K)), kb(cnt * sizeof(K));
  gen_level_keys<K><<<grid_for(cnt, BLK), BLK, 0, c.stream>>>(
      c.d_inds.as<u32>(), c.nnz, N, d_comb.as<int>(), ncomb, s, b, ka.as<K>());
  SP_CHECK_LAUNCH();
  K *sorted = nullptr;
  radix_sort<K>(ka.as<K>(), kb.as<K>(), nullptr, nullptr, cnt, s * b, c.stream,
                &sorted, nullptr);
  DevBuf flag(cnt * sizeof(u64));
  flag_heads<K><<<grid_for(cnt, BLK), BLK, 0, c.stream>>>(sorted, cnt, flag.as<u64>());
  SP_CHECK_LAUNCH();
  exclusive_scan_u64(flag.as<u64>(), flag.as<u64>(), cnt, c.stream);
  u64 last_pos = 0;
  SP_HIP(hipMemcpyAsync(&last_pos, flag.as<u64>() + cnt - 1, sizeof(u64),
                        hipMemcpyDeviceToHost, c.stream));
  K last_two[2] = {0, 0};
  SP_HIP(hipMemcpyAsync(last_two, sorted + (cnt >= 2 ? cnt - 2 : 0),
                        (cnt >= 2 ? 2 : 1) * sizeof(K), hipMemcpyDeviceToHost,
                        c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  const bool last_head = cnt == 1 || last_two[0] != last_two[1];
  const u64 L = last_pos + (last_head ? 1 : 0);
  DevLevel &lv = c.levels[l];
  lv.size = L;
  lv.keys.alloc(L * sizeof(K));
  compact_heads<K><<<grid_for(cnt, BLK), BLK, 0, c.stream>>>(
      sorted, cnt, flag.as<u64>(), lv.keys.as<K>(), nullptr, nullptr);
  SP_CHECK_LAUNCH();
  lv.level_idx.alloc(L * sizeof(u32));
  last_index<K><<<grid_for(L, BLK), BLK, 0, c.stream>>>(lv.keys.as<K>(), L, b,
                                                        lv.level_idx.as<u32>());
  SP_CHECK_LAUNCH();
  SP_HIP(hipStreamSynchronize(c.stream));
}

// Leaf level + owners (stable by nonzero order, then j descending).
template <class K>
void build_leaf_level(Ctx &c, DevBuf &own_leaf) {
  const int N = (int)c.order, l = c.depth - 1, s = N - 1, b = c.bits;
  const u64 no = c.nnz * (u64)N;
  if (no > 0xFFFFFFFFull) throw Error(SP_ERR_UNSUPPORTED, "more than 2^32 owners");
  DevBuf ka(no * sizeof(K)), kb(no * sizeof(K));
  DevBuf ra(no * sizeof(u32)), rb(no * sizeof(u32));
  gen_owner_keys<K><<<grid_for(no, BLK), BLK, 0, c.stream>>>(
      c.d_inds.as<u32>(), c.nnz, N, b, ka.as<K>(), ra.as<u32>());
  SP_CHECK_LAUNCH();
  K *sorted = nullptr;
  u32 *rec = nullptr;
  radix_sort<K>(ka.as<K>(), kb.as<K>(), ra.as<u32>(), rb.as<u32>(), no, s * b,
                c.stream, &sorted, &rec);
  DevBuf flag(no * sizeof(u64));
  flag_heads<K><<<grid_for(no, BLK), BLK, 0, c.stream>>>(sorted, no, flag.as<u64>());
  SP_CHECK_LAUNCH();
  exclusive_scan_u64(flag.as<u64>(), flag.as<u64>(), no, c.stream);
  u64 last_pos = 0;
  K last_two[2] = {0, 0};
  SP_HIP(hipMemcpyAsync(&last_pos, flag.as<u64>() + no - 1, sizeof(u64),
                        hipMemcpyDeviceToHost, c.stream));
  SP_HIP(hipMemcpyAsync(last_two, sorted + (no >= 2 ? no - 2 : 0),
                        (no >= 2 ? 2 : 1) * sizeof(K), hipMemcpyDeviceToHost,
                        c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  const bool last_head = no == 1 || last_two[0] != last_two[1];
  const u64 L = last_pos + (last_head ? 1 : 0);
  DevLevel &lv = c.levels[l];
  lv.size = L;
  lv.keys.alloc(L * sizeof(K));
  lv.level_ptr.alloc((L + 1) * sizeof(u64));
  own_leaf.alloc(no * sizeof(u64));
  compact_heads<K><<<grid_for(no, BLK), BLK, 0, c.stream>>>(
      sorted, no, flag.as<u64>(), lv.keys.as<K>(), own_leaf.as<u64>(),
      lv.level_ptr.as<u64>());
  SP_CHECK_LAUNCH();
  gather_ptr_tail<<<1, 1, 0, c.stream>>>(lv.level_ptr.as<u64>(), L, no);
  SP_CHECK_LAUNCH();
  lv.level_idx.alloc(L * sizeof(u32));
  last_index<K><<<grid_for(L, BLK), BLK, 0, c.stream>>>(lv.keys.as<K>(), L, b,
                                                        lv.level_idx.as<u32>());
  SP_CHECK_LAUNCH();
  c.n_owners = no;
  c.own_rem.alloc(no * sizeof(u32));
  c.own_val.alloc(no * sizeof(double));
  owner_fields<<<grid_for(no, BLK), BLK, 0, c.stream>>>(
      rec, no, N, c.d_inds.as<u32>(), c.d_vals.as<double>(), c.own_rem.as<u32>(),
      c.own_val.as<double>());
  SP_CHECK_LAUNCH();
  SP_HIP(hipStreamSynchronize(c.stream));
}

template <class KC, class KP>
void link_levels(Ctx &c, int l) {
  // CSR from level l (parents, KP) to level l+1 (children, KC)
  DevLevel &P = c.levels[l];
  DevLevel &C = c.levels[l + 1];
  DevBuf cfidx(((u64)c.dim + 2) * sizeof(u64));
  first_field_index<KC><<<grid_for((u64)c.dim + 2, BLK), BLK, 0, c.stream>>>(
      C.keys.as<KC>(), C.size, l + 2, c.bits, c.dim, cfidx.as<u64>());
  SP_CHECK_LAUNCH();
  P.level_ptr.alloc((P.size + 1) * sizeof(u64));
  child_ptr<KC, KP><<<grid_for(P.size + 1, BLK), BLK, 0, c.stream>>>(
      C.keys.as<KC>(), C.size, c.bits, P.keys.as<KP>(), P.size, l + 1, cfidx.as<u64>(),
      P.level_ptr.as<u64>());
  SP_CHECK_LAUNCH();
  SP_HIP(hipStreamSynchronize(c.stream));
}

// dep_list of level l+1 against level l.
template <class KC, class KP>
void exclude_level(Ctx &c, int l) {
  DevLevel &P = c.levels[l];
  DevLevel &C = c.levels[l + 1];
  DevBuf fidx(((u64)c.dim + 2) * sizeof(u64));
  first_field_index<KP><<<grid_for((u64)c.dim + 2, BLK), BLK, 0, c.stream>>>(
      P.keys.as<KP>(), P.size, l + 1, c.bits, c.dim, fidx.as<u64>());
  SP_CHECK_LAUNCH();
  const int s = l + 2;
  C.dep_list.alloc(C.size * (u64)s * sizeof(u64));
  exclude_one<KC, KP><<<grid_for(C.size * (u64)s, BLK), BLK, 0, c.stream>>>(
      C.keys.as<KC>(), C.size, s, c.bits, P.keys.as<KP>(), fidx.as<u64>(),
      C.dep_list.as<u64>());
  SP_CHECK_LAUNCH();
  SP_HIP(hipStreamSynchronize(c.stream));
}

template <class K>
void build_worklist_t(Ctx &c, DevBuf &own_leaf) {
  const int m = (int)c.order - 1;
  const u64 no = c.n_owners;
  DevBuf ka(no * sizeof(u64)), kb(no * sizeof(u64));
  DevBuf ia(no * sizeof(u32)), ib(no * sizeof(u32));
  rem_keys<<<grid_for(no, BLK), BLK, 0, c.stream>>>(c.own_rem.as<u32>(), no,
                                                    ka.as<u64>(), ia.as<u32>());
  SP_CHECK_LAUNCH();
  u64 *srem = nullptr;
  u32 *sid = nullptr;
  radix_sort<u64>(ka.as<u64>(), kb.as<u64>(), ia.as<u32>(), ib.as<u32>(), no,
                  c.bits, c.stream, &srem, &sid);
  c.w_tuple.alloc(no * m * sizeof(u32));
  c.w_coef.alloc(no * sizeof(double));
  c.w_coef32.alloc(no * sizeof(float));
  c.w_val.alloc(no * sizeof(double));
  c.w_leaf.alloc(no * sizeof(u32));
  if (c.levels[c.depth - 1].size >= (1ull << 32))
    throw Error(SP_ERR_UNSUPPORTED, "more than 2^32 leaves");
  build_worklist<K><<<grid_for(no, BLK), BLK, 0, c.stream>>>(
      sid, no, own_leaf.as<u64>(), c.own_rem.as<u32>(), c.own_val.as<double>(),
      c.levels[c.depth - 1].keys.as<K>(), m, c.bits, c.w_tuple.as<u32>(),
      c.w_coef.as<double>(), c.w_coef32.as<float>(), c.w_val.as<double>(),
      c.w_leaf.as<u32>());
  SP_CHECK_LAUNCH();
  c.w_segptr.alloc((c.dim + 1) * sizeof(u64));
  seg_bounds<<<grid_for((u64)c.dim + 1, BLK), BLK, 0, c.stream>>>(srem, no, c.dim,
                                                                  c.w_segptr.as<u64>());
  SP_CHECK_LAUNCH();
  c.h_segptr.resize(c.dim + 1);
  SP_HIP(hipMemcpyAsync(c.h_segptr.data(), c.w_segptr.p, (c.dim + 1) * sizeof(u64),
                        hipMemcpyDeviceToHost, c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
}

template <class F>
void with_key(int words, F &&f) {
  if (words == 1) f(u64{}); else f(u128{});
}

}  // namespace

namespace {

// Input checks and ||X||^2 terms (spsymtensor.cpp:114-138): block b sums the
// terms v^2 N!/prod f! of nonzeros [b*2048, (b+1)*2048) in a fixed order
// (8 consecutive terms per thread, then a binary tree over the 256 threads)
// into part[b]; a tuple with an index outside [1, dim] or out of order sets
// bit 1 or 2 of *bad.  No atomics on the sum: the result does not depend on
// scheduling.
constexpr int kNormItems = 8;

__global__ void __launch_bounds__(256)
check_norm_partial(const u32 *__restrict__ inds, const double *__restrict__ vals, u64 nnz,
                   int N, u32 dim, double *__restrict__ part, unsigned int *__restrict__ bad) {
  __shared__ double red[256];
  u64 f[20];
  f[0] = 1;
  for (int i = 1; i < 20; ++i) f[i] = f[i - 1] * (u64)i;
  const u64 base = ((u64)blockIdx.x * 256 + threadIdx.x) * kNormItems;
  double s = 0.0;
  unsigned int flags = 0;
  for (int k = 0; k < kNormItems; ++k) {
    const u64 i = base + k;
    if (i >= nnz) break;
    const u32 *t = inds + i * N;
    u64 denom = 1;
    int run = 1;
    u32 prev = t[0];
    flags |= (prev < 1 || prev > dim) ? 1u : 0u;
    for (int j = 1; j <= N; ++j) {
      const u32 x = j < N ? t[j] : 0;
      if (j < N) {
        flags |= (x < 1 || x > dim) ? 1u : 0u;
        flags |= (x < prev) ? 2u : 0u;
      }
      if (j < N && x == prev) {
        ++run;
      } else {
        denom *= f[run];
        run = 1;
      }
      prev = x;
    }
    const double v = vals[i];
    s += v * v * (double)(f[N] / denom);
  }
  if (flags) atomicOr(bad, flags);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Sum of the block partials, same fixed pattern in one block.
__global__ void __launch_bounds__(256)
sum_norm_parts(const double *__restrict__ part, u64 n, double *__restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (u64 i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

}  // namespace

void input_check_norm(Ctx &c) {
  const u64 nb = (c.nnz + 256 * kNormItems - 1) / (256 * kNormItems);
  DevBuf part((nb + 1) * sizeof(double) + 256);
  double *dpart = part.as<double>();
  double *dsum = dpart + nb;
  unsigned int *dbad = reinterpret_cast<unsigned int *>(dpart + nb + 1);
  SP_HIP(hipMemsetAsync(dbad, 0, sizeof(unsigned int), c.stream));
  check_norm_partial<<<(unsigned)nb, 256, 0, c.stream>>>(c.d_inds.as<u32>(),
                                                         c.d_vals.as<double>(), c.nnz,
                                                         (int)c.order, c.dim, dpart, dbad);
  SP_CHECK_LAUNCH();
  sum_norm_parts<<<1, 256, 0, c.stream>>>(dpart, nb, dsum);
  SP_CHECK_LAUNCH();
  double sum = 0.0;
  unsigned int bad = 0;
  SP_HIP(hipMemcpyAsync(&sum, dsum, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SP_HIP(hipMemcpyAsync(&bad, dbad, sizeof(unsigned int), hipMemcpyDeviceToHost, c.stream));
  SP_HIP(hipStreamSynchronize(c.stream));
  if (bad & 1u) throw Error(SP_ERR_ARG, "index outside [1,dim]");
  if (bad & 2u) throw Error(SP_ERR_ARG, "index tuple not sorted");
  c.norm = std::sqrt(sum);
}

// The S3TTMc path needs the leaf level, the owners and the rem-sorted work
// list only; those are built here.  The interior levels with their CSR
// links (cdtree_ensure_levels) and the exclude-one parent lists
// (cdtree_ensure_dep_list) are built on first use.
void cdtree_build(Ctx &c) {
  const int N = (int)c.order;
  c.depth = N - 1;
  int b = 1;
  while (b < 32 && (1ull << b) <= c.dim) ++b;
  c.bits = b;
  if ((N - 1) * b > 128)
    throw Error(SP_ERR_UNSUPPORTED,
                "(order-1) * bits(dim) exceeds the 128-bit sub-multiset key");
  c.key_words = key_words_for(N - 1, b);
  c.levels.clear();
  c.levels.resize(c.depth);
  c.levels_built = false;
  DevBuf own_leaf;
  if (c.key_words == 1) build_leaf_level<u64>(c, own_leaf);
  else build_leaf_level<u128>(c, own_leaf);
  if (c.key_words == 1) build_worklist_t<u64>(c, own_leaf);
  else build_worklist_t<u128>(c, own_leaf);
  c.built = true;
  c.chunk_R = 0;  // work chunks depend on the tree
}

void cdtree_ensure_levels(Ctx &c) {
  if (!c.built) throw Error(SP_ERR_STATE, "tree not built");
  if (c.levels_built) return;
  const int b = c.bits;
  for (int l = 0; l < c.depth - 1; ++l) {
    if (key_words_for(l + 1, b) == 1) build_interior_level<u64>(c, l);
    else build_interior_level<u128>(c, l);
  }
  for (int l = 0; l + 1 < c.depth; ++l) {
    const int wp = key_words_for(l + 1, b), wc = key_words_for(l + 2, b);
    if (wp == 1 && wc == 1) link_levels<u64, u64>(c, l);
    else if (wp == 1) link_levels<u128, u64>(c, l);
    else link_levels<u128, u128>(c, l);
  }
  c.levels_built = true;
}

void cdtree_ensure_dep_list(Ctx &c, int l) {
  if (l < 1 || l >= c.depth) return;
  cdtree_ensure_levels(c);
  if (c.levels[l].dep_list.p) return;
  const int wp = key_words_for(l, c.bits), wc = key_words_for(l + 1, c.bits);
  if (wp == 1 && wc == 1) exclude_level<u64, u64>(c, l - 1);
  else if (wp == 1) exclude_level<u128, u64>(c, l - 1);
  else exclude_level<u128, u128>(c, l - 1);
}

void cdtree_export_level(Ctx &c, int l, u32 *keys, u32 *level_idx, u64 *level_ptr,
                         u64 *dep_list) {
  if (l < c.depth - 1) cdtree_ensure_levels(c);
  if (dep_list) cdtree_ensure_dep_list(c, l);
  DevLevel &lv = c.levels[l];
  const int s = l + 1;
  const u64 L = lv.size;
  if (keys) {
    const int w = key_words_for(s, c.bits);
    std::vector<u64> raw(L * w);
    SP_HIP(hipMemcpy(raw.data(), lv.keys.p, L * w * sizeof(u64), hipMemcpyDeviceToHost));
    const u128 mask = (u128)((1ull << c.bits) - 1);
    for (u64 j = 0; j < L; ++j) {
      u128 k = 0;
      if (w == 1) k = raw[j];
      else std::memcpy(&k, &raw[2 * j], sizeof(u128));
      for (int i = 0; i < s; ++i)
        keys[j * s + i] = (u32)((k >> (c.bits * (s - 1 - i))) & mask);
    }
  }
  if (level_idx)
    SP_HIP(hipMemcpy(level_idx, lv.level_idx.p, L * sizeof(u32), hipMemcpyDeviceToHost));
  if (level_ptr)
    SP_HIP(hipMemcpy(level_ptr, lv.level_ptr.p, (L + 1) * sizeof(u64),
                     hipMemcpyDeviceToHost));
  if (dep_list && l >= 1)
    SP_HIP(hipMemcpy(dep_list, lv.dep_list.p, L * s * sizeof(u64),
                     hipMemcpyDeviceToHost));
}

void cdtree_export_owners(Ctx &c, u32 *rem, double *val) {
  if (rem)
    SP_HIP(hipMemcpy(rem, c.own_rem.p, c.n_owners * sizeof(u32), hipMemcpyDeviceToHost));
  if (val)
    SP_HIP(hipMemcpy(val, c.own_val.p, c.n_owners * sizeof(double),
                     hipMemcpyDeviceToHost));
}

}  // namespace spamd
