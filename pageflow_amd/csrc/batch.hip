// Batched equi-joins: many independent inner joins of uneven size run as
// one pipeline of persistent kernels that pull work items from device
// queues, so a batch of 256 joins costs a handful of launches and no host
// wait; the steady state is captured once as a hipGraph and replayed.
//
// Each join j: int32 build keys (unique) and int32 probe keys, stored back
// to back in two buffers (offsets b_off / p_off).  Output: join j's matched
// (probe row, build row) pairs, join-local int64 row ids, in probe-row
// order, in out_probe/out_build at p_off[j] (the upper bound of a unique
// build's matches), their count in counts[j].
//
//   bj_build    items = (join, 4096 build rows): insert (key << 32 | row)
//               into the join's slice of one table arena with a 64-bit CAS
//               (empty = all ones, never a valid slot: a row id is < 2^31);
//               a repeated key keeps its lowest row (atomicMin)
//   bj_probe    items = (join, probe chunk): each probe row's build row or
//               -1 into partner, the item's matches into item_cnt
//   scan        item_cnt -> item_pos (items are join-major, rows ascending)
//   bj_compact  items again: ballot-ranked stable compaction of matches
//   bj_counts   per-join match counts
//
// Tunables per size class of the build (small < 32K rows, medium < 512K,
// large): the table's slots per build row (2 or 4, rounded to a power of
// two), the probe chunk (4096 or 16384 rows), and the workgroup size of the
// probe and compaction kernels (256 or 512).  pf_batch_plan_create takes
// them (the autotuner in pageflow_amd/batch.py measures and picks them).
//
// The plain baseline (mode 1) runs the same kernels one join at a time,
// with the reference's one host wait per join (kernel_utils.cuh:37-40).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "pf_common.hpp"
#include "../../include/pageflow_join.h"

namespace pf {
namespace {

constexpr uint64_t kEmpty = ~0ULL;
constexpr int BUILD_ROWS = 4096;
constexpr int NCLASS = 3;

struct Tune {
  int load;   // slots per build row (2 or 4)
  int chunk;  // probe rows an item
  int nt;     // probe / compact workgroup size (256 or 512)
};

int class_of(int64_t nb) { return nb < 32768 ? 0 : (nb < 524288 ? 1 : 2); }

struct Plan {
  int nj = 0;
  std::vector<int64_t> nb, np, b_off, p_off, t_off, cap;
  std::vector<int> cls;
  Tune tune[NCLASS];
  // items: build items all joins; probe items grouped by workgroup size
  int64_t n_bitems = 0;
  int64_t n_pitems[2] = {0, 0};      // [0]: nt 256, [1]: nt 512
  int64_t pitem_base[2] = {0, 0};    // first item of each group in the item arrays
  int64_t n_pitems_all = 0;
  int64_t total_slots = 0, total_b = 0, total_p = 0;
  std::vector<int64_t> first_item, last_item;  // per join, in the probe item arrays
  // device
  void* dmem = nullptr;
  size_t dbytes = 0;
  int64_t *d_b_off, *d_p_off, *d_t_off;
  uint32_t* d_mask;
  int32_t* d_bitem_join;
  int64_t* d_bitem_start;
  int32_t *d_pitem_join, *d_pitem_len;
  int64_t* d_pitem_start;
  int64_t *d_item_cnt, *d_item_pos;
  int64_t *d_first, *d_last;
  int32_t* d_partner;
  uint64_t* d_table;
  int32_t* d_queue;  // [0] build, [1] probe 256, [2] probe 512, [3] compact 256, [4] compact 512
  void* d_scan;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int persist = 1024;  // workgroups of a persistent launch
};

// A work item taken from the queue, as a wave-uniform (scalar) value.  The
// persistent loops below exit on it and hold barriers; the compiler must see
// the exit as uniform, otherwise it may nest the loop so that the barriers
// and the read of the next item run in an inner loop that the lane fetching
// items (thread 0) never re-enters, and the other lanes spin on a stale item.
__device__ __forceinline__ int64_t uniform_item(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ uint32_t bslot(int32_t key, uint32_t mask) {
  return (uint32_t)mix64((uint64_t)(int64_t)key) & mask;
}

__global__ void bj_fill_tables(uint64_t* table, int64_t n, int32_t* queue, int nq) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t i = t0; i < n; i += stride) table[i] = kEmpty;
  if (t0 < nq) queue[t0] = 0;
}

// Build items: (join, first row) with BUILD_ROWS rows (fewer at a join's end).
__global__ __launch_bounds__(256) void bj_build(const int32_t* __restrict__ bkeys,
                                                const int64_t* __restrict__ b_off,
                                                const int64_t* __restrict__ t_off,
                                                const uint32_t* __restrict__ mask,
                                                const int32_t* __restrict__ item_join,
                                                const int64_t* __restrict__ item_start,
                                                int64_t n_items, int64_t item_lo,
                                                uint64_t* table, int32_t* queue) {
  __shared__ int64_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = item_lo + atomicAdd(queue, 1);
    __syncthreads();
    const int64_t item = uniform_item(s_item);
    __syncthreads();
    if (item >= n_items) return;
    const int j = item_join[item];
    const int64_t r0 = item_start[item];
    const int64_t nrows = b_off[j + 1] - b_off[j];
    const int64_t r1 = r0 + BUILD_ROWS < nrows ? r0 + BUILD_ROWS : nrows;
    uint64_t* t = table + t_off[j];
    const uint32_t m = mask[j];
    // U rows a thread with their CASes in flight together (one at a time is
    // latency-bound).  No read before the CAS: in a fresh table of unique
    // keys the first slot is usually empty and the CAS returns what a read
    // would have.
    constexpr int U = 4;
    for (int64_t rr = r0 + threadIdx.x; rr < r1; rr += 256 * U) {
      int32_t k[U];
      uint32_t s[U];
      bool todo[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t r = rr + u * 256;
        todo[u] = r < r1;
        k[u] = todo[u] ? bkeys[b_off[j] + r] : 0;
        s[u] = bslot(k[u], m);
      }
      bool any = true;
      while (any) {
        uint64_t prev[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint64_t v = ((uint64_t)(uint32_t)k[u] << 32) | (uint32_t)(rr + u * 256);
          prev[u] = todo[u] ? atomicCAS((unsigned long long*)&t[s[u]], (unsigned long long)kEmpty,
                                        (unsigned long long)v)
                            : kEmpty;
        }
        any = false;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (!todo[u]) continue;
          const uint64_t v = ((uint64_t)(uint32_t)k[u] << 32) | (uint32_t)(rr + u * 256);
          if (prev[u] == kEmpty) {
            todo[u] = false;
          } else if ((uint32_t)(prev[u] >> 32) == (uint32_t)k[u]) {
            atomicMin((unsigned long long*)&t[s[u]], (unsigned long long)v);
            todo[u] = false;
          } else {
            s[u] = (s[u] + 1) & m;
            any = true;
          }
        }
      }
    }
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void bj_probe(const int32_t* __restrict__ pkeys,
                                               const int64_t* __restrict__ p_off,
                                               const int64_t* __restrict__ t_off,
                                               const uint32_t* __restrict__ mask,
                                               const int32_t* __restrict__ item_join,
                                               const int64_t* __restrict__ item_start,
                                               const int32_t* __restrict__ item_len,
                                               int64_t n_items, int64_t item_lo,
                                               const uint64_t* __restrict__ table,
                                               int32_t* __restrict__ partner,
                                               int64_t* __restrict__ item_cnt, int32_t* queue) {
  __shared__ int64_t s_item;
  __shared__ int red[NT / kWave];
  // one thread-0 region an iteration (the previous item's count, then the
  // next item), as in the other persistent loops
  int64_t prev = -1;
  int prev_tot = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      if (prev >= 0) item_cnt[prev] = prev_tot;
      s_item = item_lo + atomicAdd(queue, 1);
    }
    __syncthreads();
    const int64_t item = uniform_item(s_item);
    __syncthreads();
    if (item >= item_lo + n_items) return;
    const int j = item_join[item];
    const int64_t base = p_off[j] + item_start[item];
    const int len = item_len[item];
    const uint64_t* t = table + t_off[j];
    const uint32_t m = mask[j];
    int c = 0;
    // U probe rows a thread with their first-slot reads in flight together;
    // the rare longer probe sequences are resolved afterwards
    constexpr int U = 8;
    for (int r0 = threadIdx.x; r0 < len; r0 += NT * U) {
      int32_t k[U];
      uint32_t s[U];
      uint64_t v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = r0 + u * NT;
        k[u] = r < len ? pkeys[base + r] : 0;
        s[u] = bslot(k[u], m);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = r0 + u * NT < len ? t[s[u]] : kEmpty;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = r0 + u * NT;
        if (r >= len) break;
        while (v[u] != kEmpty && (uint32_t)(v[u] >> 32) != (uint32_t)k[u]) {
          s[u] = (s[u] + 1) & m;
          v[u] = t[s[u]];
        }
        const int32_t hit = v[u] == kEmpty ? -1 : (int32_t)(uint32_t)v[u];
        partner[base + r] = hit;
        c += hit >= 0;
      }
    }
    c = wave_sum(c);
    if (lane_id() == 0) red[threadIdx.x / kWave] = c;
    __syncthreads();
    prev_tot = 0;
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) prev_tot += red[w];
    prev = item;
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void bj_compact(const int64_t* __restrict__ p_off,
                                                 const int32_t* __restrict__ item_join,
                                                 const int64_t* __restrict__ item_start,
                                                 const int32_t* __restrict__ item_len,
                                                 int64_t n_items, int64_t item_lo,
                                                 const int64_t* __restrict__ first,
                                                 const int32_t* __restrict__ partner,
                                                 const int64_t* __restrict__ item_pos,
                                                 int64_t* __restrict__ out_probe,
                                                 int64_t* __restrict__ out_build, int32_t* queue) {
  __shared__ int64_t s_item;
  __shared__ int cnt[NT / kWave];
  __shared__ int64_t s_base;
  constexpr int W = NT / kWave;
  for (;;) {
    if (threadIdx.x == 0) s_item = item_lo + atomicAdd(queue, 1);
    __syncthreads();
    const int64_t item = uniform_item(s_item);
    __syncthreads();
    if (item >= item_lo + n_items) return;
    const int j = item_join[item];
    const int64_t r00 = item_start[item];
    const int64_t base = p_off[j] + r00;
    const int len = item_len[item];
    int64_t o = p_off[j] + item_pos[item] - item_pos[first[j]];
    for (int r0 = 0; r0 < len; r0 += NT) {
      const int r = r0 + threadIdx.x;
      const int32_t pr = r < len ? partner[base + r] : -1;
      const uint64_t bal = ballot(pr >= 0);
      const int w = threadIdx.x / kWave;
      if (lane_id() == 0) cnt[w] = __popcll(bal);
      __syncthreads();
      int pre = 0, tot = 0;
#pragma unroll
      for (int q = 0; q < W; ++q) {
        pre += q < w ? cnt[q] : 0;
        tot += cnt[q];
      }
      if (pr >= 0) {
        const int64_t pos = o + pre + mbcnt(bal);
        out_probe[pos] = r00 + r;
        out_build[pos] = pr;
      }
      o += tot;
      __syncthreads();
    }
    (void)s_base;
  }
}

__global__ void bj_counts(const int64_t* __restrict__ item_pos, const int64_t* __restrict__ first,
                          const int64_t* __restrict__ last, int nj, int64_t* __restrict__ counts) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nj) counts[j] = first[j] < 0 ? 0 : item_pos[last[j] + 1] - item_pos[first[j]];
}

int64_t pow2_at_least(int64_t x) {
  int64_t c = 16;
  while (c < x) c <<= 1;
  return c;
}

int plan_build(Plan& P, int nj, const int64_t* nb, const int64_t* np, const int32_t* tune) {
  P.nj = nj;
  for (int c = 0; c < NCLASS; ++c) {
    P.tune[c] = Tune{2, c == 0 ? 4096 : 16384, 256};
    if (tune) {
      P.tune[c].load = tune[c * 3 + 0];
      P.tune[c].chunk = tune[c * 3 + 1];
      P.tune[c].nt = tune[c * 3 + 2];
    }
    const Tune& t = P.tune[c];
    if ((t.load != 2 && t.load != 4) || t.chunk < 256 || t.chunk > (1 << 20) ||
        (t.nt != 256 && t.nt != 512))
      return fail("batch plan: tunables are load 2|4, chunk in [256, 2^20], nt 256|512");
  }
  P.nb.assign(nb, nb + nj);
  P.np.assign(np, np + nj);
  P.b_off.assign(nj + 1, 0);
  P.p_off.assign(nj + 1, 0);
  P.t_off.assign(nj + 1, 0);
  P.cap.assign(nj, 0);
  P.cls.assign(nj, 0);
  for (int j = 0; j < nj; ++j) {
    if (nb[j] < 0 || np[j] < 0 || nb[j] >= (1LL << 31) || np[j] >= (1LL << 31))
      return fail("batch plan: a join side takes fewer than 2^31 rows");
    P.cls[j] = class_of(nb[j]);
    P.cap[j] = pow2_at_least(P.tune[P.cls[j]].load * std::max<int64_t>(nb[j], 1));
    P.b_off[j + 1] = P.b_off[j] + nb[j];
    P.p_off[j + 1] = P.p_off[j] + np[j];
    P.t_off[j + 1] = P.t_off[j] + P.cap[j];
  }
  P.total_b = P.b_off[nj];
  P.total_p = P.p_off[nj];
  P.total_slots = P.t_off[nj];
  return 0;
}

}  // namespace
}  // namespace pf

extern "C" {

void* pf_batch_plan_create(int32_t njoins, const int64_t* nb, const int64_t* np,
                           const int32_t* tune) {
  using namespace pf;
  if (njoins < 1) {
    fail("batch plan: at least one join");
    return nullptr;
  }
  Plan* P = new Plan();
  if (plan_build(*P, njoins, nb, np, tune)) {
    delete P;
    return nullptr;
  }
  // host item lists
  std::vector<int32_t> bj, pj[2], pl[2];
  std::vector<int64_t> bs, ps[2];
  for (int j = 0; j < njoins; ++j)
    for (int64_t r = 0; r < P->nb[j]; r += BUILD_ROWS) {
      bj.push_back(j);
      bs.push_back(r);
    }
  for (int j = 0; j < njoins; ++j) {
    const Tune& t = P->tune[P->cls[j]];
    const int g = t.nt == 512 ? 1 : 0;
    if (P->nb[j] == 0) continue;  // no match possible
    for (int64_t r = 0; r < P->np[j]; r += t.chunk) {
      pj[g].push_back(j);
      ps[g].push_back(r);
      pl[g].push_back((int32_t)std::min<int64_t>(t.chunk, P->np[j] - r));
    }
  }
  // probe items: group 0 then group 1; per-join first/last item
  std::vector<int32_t> ij, il;
  std::vector<int64_t> is;
  for (int g = 0; g < 2; ++g) {
    P->pitem_base[g] = (int64_t)ij.size();
    P->n_pitems[g] = (int64_t)pj[g].size();
    ij.insert(ij.end(), pj[g].begin(), pj[g].end());
    is.insert(is.end(), ps[g].begin(), ps[g].end());
    il.insert(il.end(), pl[g].begin(), pl[g].end());
  }
  P->n_pitems_all = (int64_t)ij.size();
  P->n_bitems = (int64_t)bj.size();
  P->first_item.assign(njoins, -1);
  P->last_item.assign(njoins, -1);
  for (int64_t i = 0; i < P->n_pitems_all; ++i) {
    const int j = ij[i];
    if (P->first_item[j] < 0) P->first_item[j] = i;
    P->last_item[j] = i;
  }
  // device memory: one allocation
  const int64_t nbi = std::max<int64_t>(P->n_bitems, 1), npi = std::max<int64_t>(P->n_pitems_all, 1);
  Carve cv(nullptr);
  auto layout = [&](Carve& c) {
    P->d_b_off = c.take<int64_t>(njoins + 1);
    P->d_p_off = c.take<int64_t>(njoins + 1);
    P->d_t_off = c.take<int64_t>(njoins + 1);
    P->d_mask = c.take<uint32_t>(njoins);
    P->d_bitem_join = c.take<int32_t>(nbi);
    P->d_bitem_start = c.take<int64_t>(nbi);
    P->d_pitem_join = c.take<int32_t>(npi);
    P->d_pitem_len = c.take<int32_t>(npi);
    P->d_pitem_start = c.take<int64_t>(npi);
    P->d_item_cnt = c.take<int64_t>(npi + 1);
    P->d_item_pos = c.take<int64_t>(npi + 1);
    P->d_first = c.take<int64_t>(njoins);
    P->d_last = c.take<int64_t>(njoins);
    P->d_partner = c.take<int32_t>(std::max<int64_t>(P->total_p, 1));
    P->d_table = c.take<uint64_t>(std::max<int64_t>(P->total_slots, 1));
    P->d_queue = c.take<int32_t>(8);
    P->d_scan = c.take<char>(scan_i64_scratch(npi + 1));
  };
  layout(cv);
  P->dbytes = cv.bytes();
  if (hipMalloc(&P->dmem, P->dbytes) != hipSuccess) {
    fail("batch plan: cannot allocate " + std::to_string(P->dbytes) + " device bytes");
    delete P;
    return nullptr;
  }
  Carve real(P->dmem);
  layout(real);
  std::vector<uint32_t> mask(njoins);
  for (int j = 0; j < njoins; ++j) mask[j] = (uint32_t)(P->cap[j] - 1);
  auto up = [&](void* d, const void* h, size_t bytes) {
    return bytes == 0 || hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  bool okc = up(P->d_b_off, P->b_off.data(), (njoins + 1) * 8) &&
             up(P->d_p_off, P->p_off.data(), (njoins + 1) * 8) &&
             up(P->d_t_off, P->t_off.data(), (njoins + 1) * 8) &&
             up(P->d_mask, mask.data(), njoins * 4) &&
             up(P->d_bitem_join, bj.data(), bj.size() * 4) &&
             up(P->d_bitem_start, bs.data(), bs.size() * 8) &&
             up(P->d_pitem_join, ij.data(), ij.size() * 4) &&
             up(P->d_pitem_len, il.data(), il.size() * 4) &&
             up(P->d_pitem_start, is.data(), is.size() * 8) &&
             up(P->d_first, P->first_item.data(), njoins * 8) &&
             up(P->d_last, P->last_item.data(), njoins * 8);
  if (!okc) {
    fail("batch plan: upload failed");
    (void)hipFree(P->dmem);
    delete P;
    return nullptr;
  }
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  P->persist = std::max(64, ncu * 4);
  ok();
  return P;
}

int pf_batch_plan_destroy(void* plan) {
  using namespace pf;
  Plan* P = (Plan*)plan;
  if (!P) return ok();
  if (P->exec) (void)hipGraphExecDestroy(P->exec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  if (P->dmem) (void)hipFree(P->dmem);
  delete P;
  return ok();
}

int64_t pf_batch_plan_info(void* plan, int32_t what) {
  pf::Plan* P = (pf::Plan*)plan;
  if (!P) return -1;
  switch (what) {
    case 0: return P->total_b;
    case 1: return P->total_p;
    case 2: return P->total_slots;
    case 3: return P->n_bitems;
    case 4: return P->n_pitems_all;
    case 5: return (int64_t)P->dbytes;
    default: return -1;
  }
}

}  // extern "C"

namespace pf {
namespace {

int run_batched(Plan& P, const int32_t* bkeys, const int32_t* pkeys, int64_t* out_probe,
                int64_t* out_build, int64_t* counts, hipStream_t st) {
  {
    KTimer _kt("bj_fill", st);
    bj_fill_tables<<<2048, 256, 0, st>>>(P.d_table, P.total_slots, P.d_queue, 8);
  }
  if (P.n_bitems > 0) {
    KTimer _kt("bj_build", st);
    bj_build<<<(unsigned)std::min<int64_t>(P.persist, P.n_bitems), 256, 0, st>>>(
        bkeys, P.d_b_off, P.d_t_off, P.d_mask, P.d_bitem_join, P.d_bitem_start, P.n_bitems, 0,
        P.d_table, P.d_queue + 0);
  }
  for (int g = 0; g < 2; ++g) {
    if (P.n_pitems[g] == 0) continue;
    KTimer _kt("bj_probe", st);
    const unsigned grid = (unsigned)std::min<int64_t>(P.persist, P.n_pitems[g]);
    if (g == 0)
      bj_probe<256><<<grid, 256, 0, st>>>(pkeys, P.d_p_off, P.d_t_off, P.d_mask, P.d_pitem_join,
                                          P.d_pitem_start, P.d_pitem_len, P.n_pitems[g],
                                          P.pitem_base[g], P.d_table, P.d_partner,
                                          P.d_item_cnt, P.d_queue + 1);
    else
      bj_probe<512><<<grid, 512, 0, st>>>(pkeys, P.d_p_off, P.d_t_off, P.d_mask, P.d_pitem_join,
                                          P.d_pitem_start, P.d_pitem_len, P.n_pitems[g],
                                          P.pitem_base[g], P.d_table, P.d_partner,
                                          P.d_item_cnt, P.d_queue + 2);
  }
  if (launch_check("batch join")) return 1;
  if (scan_i64_excl_from(P.d_item_cnt, P.d_item_pos, P.n_pitems_all, P.d_item_pos +
                         std::max<int64_t>(P.n_pitems_all, 0), P.d_scan, st))
    return 1;
  for (int g = 0; g < 2; ++g) {
    if (P.n_pitems[g] == 0) continue;
    KTimer _kt("bj_compact", st);
    const unsigned grid = (unsigned)std::min<int64_t>(P.persist, P.n_pitems[g]);
    if (g == 0)
      bj_compact<256><<<grid, 256, 0, st>>>(P.d_p_off, P.d_pitem_join, P.d_pitem_start,
                                            P.d_pitem_len, P.n_pitems[g], P.pitem_base[g],
                                            P.d_first, P.d_partner, P.d_item_pos, out_probe,
                                            out_build, P.d_queue + 3);
    else
      bj_compact<512><<<grid, 512, 0, st>>>(P.d_p_off, P.d_pitem_join, P.d_pitem_start,
                                            P.d_pitem_len, P.n_pitems[g], P.pitem_base[g],
                                            P.d_first, P.d_partner, P.d_item_pos, out_probe,
                                            out_build, P.d_queue + 4);
  }
  bj_counts<<<grid_for(P.nj, 256), 256, 0, st>>>(P.d_item_pos, P.d_first, P.d_last, P.nj,
                                                 counts);
  return launch_check("batch join");
}

// One join at a time through the same kernels, waiting for each join's
// count as a per-call entry point does.
int run_per_join(Plan& P, const int32_t* bkeys, const int32_t* pkeys, int64_t* out_probe,
                 int64_t* out_build, int64_t* counts, hipStream_t st) {
  // item ranges per join in the build list
  std::vector<int64_t> bfirst(P.nj + 1, 0);
  for (int j = 0; j < P.nj; ++j)
    bfirst[j + 1] = bfirst[j] + (P.nb[j] + BUILD_ROWS - 1) / BUILD_ROWS;
  for (int j = 0; j < P.nj; ++j) {
    bj_fill_tables<<<grid_for(P.cap[j], 256) < 2048 ? grid_for(P.cap[j], 256) : 2048, 256, 0,
                     st>>>(P.d_table + P.t_off[j], P.cap[j], P.d_queue, 8);
    const int64_t nbi = bfirst[j + 1] - bfirst[j];
    if (nbi > 0)
      bj_build<<<(unsigned)std::min<int64_t>(P.persist, nbi), 256, 0, st>>>(
          bkeys, P.d_b_off, P.d_t_off, P.d_mask, P.d_bitem_join, P.d_bitem_start,
          bfirst[j + 1], bfirst[j], P.d_table, P.d_queue + 0);
    const int64_t f = P.first_item[j], l = P.last_item[j];
    if (f >= 0) {
      const int64_t n = l - f + 1;
      const unsigned grid = (unsigned)std::min<int64_t>(P.persist, n);
      const bool big = P.tune[P.cls[j]].nt == 512;
      if (!big)
        bj_probe<256><<<grid, 256, 0, st>>>(pkeys, P.d_p_off, P.d_t_off, P.d_mask,
                                            P.d_pitem_join, P.d_pitem_start, P.d_pitem_len, n, f,
                                            P.d_table, P.d_partner, P.d_item_cnt,
                                            P.d_queue + 1);
      else
        bj_probe<512><<<grid, 512, 0, st>>>(pkeys, P.d_p_off, P.d_t_off, P.d_mask,
                                            P.d_pitem_join, P.d_pitem_start, P.d_pitem_len, n, f,
                                            P.d_table, P.d_partner, P.d_item_cnt,
                                            P.d_queue + 1);
      // this join's items scanned on their own; positions relative to f
      if (scan_i64_excl_from(P.d_item_cnt + f, P.d_item_pos + f, n, P.d_item_pos + f + n,
                             P.d_scan, st))
        return 1;
      if (!big)
        bj_compact<256><<<grid, 256, 0, st>>>(P.d_p_off, P.d_pitem_join, P.d_pitem_start,
                                              P.d_pitem_len, n, f, P.d_first, P.d_partner,
                                              P.d_item_pos, out_probe, out_build,
                                              P.d_queue + 3);
      else
        bj_compact<512><<<grid, 512, 0, st>>>(P.d_p_off, P.d_pitem_join, P.d_pitem_start,
                                              P.d_pitem_len, n, f, P.d_first, P.d_partner,
                                              P.d_item_pos, out_probe, out_build,
                                              P.d_queue + 3);
    }
    bj_counts<<<1, 64, 0, st>>>(P.d_item_pos, P.d_first + j, P.d_last + j, 1, counts + j);
    if (launch_check("batch join (per join)")) return 1;
    if (wait_stream(st, "batch join (per join)")) return 1;
  }
  return 0;
}

}  // namespace
}  // namespace pf

extern "C" {

int pf_batch_join(void* plan, const int32_t* bkeys, const int32_t* pkeys, int64_t* out_probe,
                  int64_t* out_build, int64_t* counts, int32_t mode, void* stream) {
  using namespace pf;
  Plan* P = (Plan*)plan;
  if (!P) return fail("batch join: no plan");
  hipStream_t st = (hipStream_t)stream;
  int rc = mode == 1 ? run_per_join(*P, bkeys, pkeys, out_probe, out_build, counts, st)
                     : run_batched(*P, bkeys, pkeys, out_probe, out_build, counts, st);
  if (rc) return rc;
  return ok();
}

int pf_batch_capture(void* plan, const int32_t* bkeys, const int32_t* pkeys, int64_t* out_probe,
                     int64_t* out_build, int64_t* counts, void* stream) {
  using namespace pf;
  Plan* P = (Plan*)plan;
  if (!P) return fail("batch capture: no plan");
  hipStream_t st = (hipStream_t)stream;
  if (P->exec) {
    (void)hipGraphExecDestroy(P->exec);
    P->exec = nullptr;
  }
  if (P->graph) {
    (void)hipGraphDestroy(P->graph);
    P->graph = nullptr;
  }
  if (hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "batch capture"))
    return 1;
  const int rc = run_batched(*P, bkeys, pkeys, out_probe, out_build, counts, st);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (hip_check(e, "batch capture")) return 1;
  P->graph = g;
  if (hip_check(hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0), "batch capture"))
    return 1;
  return ok();
}

int pf_batch_replay(void* plan, void* stream) {
  using namespace pf;
  Plan* P = (Plan*)plan;
  if (!P || !P->exec) return fail("batch replay: nothing captured");
  if (hip_check(hipGraphLaunch(P->exec, (hipStream_t)stream), "batch replay")) return 1;
  return ok();
}

}  // extern "C"
