// Stable LSD radix sort of (bin key, point index) pairs with digits of up to
// 11 bits (keys below 2^11, such as the bins of the 3D band grids, take one
// pass), 4096-key tiles.
//
// Per digit pass:
//   k_rs_upsweep   digit counts of every 4096-key tile (LDS histogram)
//   k_rs_scan_seg  exclusive scan of the digit-major count table in
//                  segments of 4096 entries (one workgroup each, coalesced)
//   k_rs_scan_top  exclusive scan of the segment totals (one workgroup)
//   k_rs_scatter   every tile moves its pairs to their places in index
//                  order: 256 keys per round, each wavefront ranks its lanes
//                  among equal digits with one ballot per digit bit (lanes
//                  with the same digit = the AND of the matching bit masks),
//                  and per-wavefront counts in LDS order the wavefronts.
// Equal keys keep their input order, so the result equals any stable sort's;
// no launch records a memory fill, so the sort can sit in a captured graph.
// An optional payload (a record and a cell word per element, indexed by the
// element's value) is moved by the last pass's scatter: with one pass and
// values 0..n-1 its reads are coalesced.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sort.hpp"

namespace fresco_hip {

namespace {

constexpr int RS_T = 256;                  // threads per tile workgroup
constexpr int RS_ROUNDS = 16;              // rounds of 256 keys per tile
constexpr int RS_TILE = RS_T * RS_ROUNDS;  // keys per tile
constexpr int RS_MAXB = 2048;              // buckets of an 11-bit digit
constexpr int RS_SEG = 4096;               // scan segment

// Workgroups go round-robin to the 8 XCDs; this maps workgroup b of nt to a
// tile so that every XCD takes a contiguous range of tiles: neighbouring
// tiles write neighbouring places of a run (shared cache lines), which then
// fill in one L2.
__device__ __forceinline__ int xcd_tile(int b, int nt) {
  const int xq = nt >> 3, xr = nt & 7, xc = b & 7;
  return xc * xq + (xc < xr ? xc : xr) + (b >> 3);
}

__global__ void __launch_bounds__(RS_T)
    k_rs_upsweep(const uint32_t* __restrict__ keys, int64_t n, int shift, uint32_t dmask, int nbuck,
                 uint32_t* __restrict__ cnt, int nblk) {
  __shared__ uint32_t h[RS_MAXB];
  const int t = threadIdx.x;
  for (int d = t; d < nbuck; d += RS_T) h[d] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const int64_t e = base + (int64_t)r * RS_T + t;
    if (e < n) atomicAdd(&h[(keys[e] >> shift) & dmask], 1u);
  }
  __syncthreads();
  for (int d = t; d < nbuck; d += RS_T) cnt[(int64_t)d * nblk + blockIdx.x] = h[d];
}

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
  // 256 threads: Hillis-Steele in LDS
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < RS_T; off <<= 1) {
    const uint32_t a = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  total = sh[RS_T - 1];
  const uint32_t incl = sh[t];
  __syncthreads();
  return incl - v;
}

// segment s = entries [s * RS_SEG, (s+1) * RS_SEG): exclusive scan in place,
// the segment total to tops[s]; thread t owns 16 consecutive entries
__global__ void __launch_bounds__(RS_T) k_rs_scan_seg(uint32_t* __restrict__ c, int64_t m, uint32_t* __restrict__ tops) {
  __shared__ uint32_t sh[RS_T];
  constexpr int PER = RS_SEG / RS_T;
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * RS_SEG + (int64_t)t * PER;
  uint32_t v[PER];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    v[i] = base + i < m ? c[base + i] : 0u;
    s += v[i];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(s, sh, total);
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    if (base + i < m) c[base + i] = run;
    run += v[i];
  }
  if (t == 0) tops[blockIdx.x] = total;
}

// exclusive scan of the segment totals (nseg entries), one workgroup
__global__ void __launch_bounds__(RS_T) k_rs_scan_top(uint32_t* __restrict__ tops, int nseg) {
  __shared__ uint32_t sh[RS_T];
  const int t = threadIdx.x;
  const int per = (nseg + RS_T - 1) / RS_T;
  const int lo = t * per < nseg ? t * per : nseg;
  const int hi = lo + per < nseg ? lo + per : nseg;
  uint32_t s = 0;
  for (int i = lo; i < hi; ++i) s += tops[i];
  uint32_t total;
  uint32_t run = block_exclusive_scan(s, sh, total);
  for (int i = lo; i < hi; ++i) {
    const uint32_t a = tops[i];
    tops[i] = run;
    run += a;
  }
}

__global__ void __launch_bounds__(RS_T)
    k_rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, int64_t n, int shift,
                 uint32_t dmask, int dbits, int nbuck, const uint32_t* __restrict__ offs,
                 const uint32_t* __restrict__ tops, int nblk, uint32_t* __restrict__ kout,
                 uint32_t* __restrict__ vout, const uint4* __restrict__ pin, uint4* __restrict__ pout, int pq,
                 const uint32_t* __restrict__ cin, uint32_t* __restrict__ cout) {
  __shared__ uint32_t run[RS_MAXB];
  __shared__ uint32_t wc[RS_T / 64][RS_MAXB];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wv = t >> 6;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  for (int d = t; d < nbuck; d += RS_T) {
    const int64_t i = (int64_t)d * nblk + tile;
    run[d] = offs[i] + tops[i / RS_SEG];
  }
  for (int d = t; d < nbuck; d += RS_T) {
#pragma unroll
    for (int w = 0; w < RS_T / 64; ++w) wc[w][d] = 0;
  }
  __syncthreads();
  const uint64_t lt = (1ull << lane) - 1ull;
  const int64_t base = (int64_t)tile * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const int64_t e = base + (int64_t)r * RS_T + t;
    const bool valid = e < n;
    const uint32_t k = valid ? kin[e] : 0u;
    const uint32_t dg = (k >> shift) & dmask;
    uint64_t same = __ballot(valid);
    for (int b = 0; b < dbits; ++b) {
      const bool bit = (dg >> b) & 1u;
      const uint64_t m = __ballot(bit);
      same &= bit ? m : ~m;
    }
    const int rank = __popcll(same & lt);
    if (valid && rank == 0) wc[wv][dg] = (uint32_t)__popcll(same);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[dg] + (uint32_t)rank;
      for (int w = 0; w < wv; ++w) pos += wc[w][dg];
      const uint32_t v = vin[e];
      kout[pos] = k;
      vout[pos] = v;
      if (pout) {
        for (int q = 0; q < pq; ++q) pout[(int64_t)pos * pq + q] = pin[(int64_t)v * pq + q];
        cout[pos] = cin[v];
      }
    }
    __syncthreads();
    // only the digits of this round move: their leaders advance the running
    // offsets (integer LDS adds: the order does not matter) and clear their
    // counts, so a round costs O(keys), not O(buckets)
    if (valid && rank == 0) {
      atomicAdd(&run[dg], wc[wv][dg]);
      wc[wv][dg] = 0;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// segmented single pass: tiles never straddle segments; counts are
// tile-major cnt[t * nbuck + d]
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(RS_T)
    k_rs_seg_upsweep(const uint32_t* __restrict__ keys, const int64_t* __restrict__ tb,
                     const int64_t* __restrict__ te, uint32_t nbuck, uint32_t* __restrict__ cnt) {
  __shared__ uint32_t h[RS_MAXB];
  const int t = threadIdx.x;
  for (int d = t; d < (int)nbuck; d += RS_T) h[d] = 0;
  __syncthreads();
  const int64_t beg = tb[blockIdx.x], end = te[blockIdx.x];
  for (int64_t e = beg + t; e < end; e += RS_T) atomicAdd(&h[keys[e] % nbuck], 1u);
  __syncthreads();
  for (int d = t; d < (int)nbuck; d += RS_T) cnt[(int64_t)blockIdx.x * nbuck + d] = h[d];
}

// one workgroup per segment: exclusive offsets over (digit, tile) in
// digit-major order, plus the segment's start
__global__ void __launch_bounds__(RS_T)
    k_rs_seg_scan(uint32_t* __restrict__ cnt, const int* __restrict__ tile0, const int64_t* __restrict__ sstart,
                  int nbuck) {
  __shared__ uint32_t sh[RS_T];
  const int sg = blockIdx.x;
  const int t0 = tile0[sg], nt = tile0[sg + 1] - t0;
  const int64_t m = (int64_t)nt * nbuck;  // entries, digit-major: e = d * nt + (t - t0)
  const int64_t per = (m + RS_T - 1) / RS_T;
  const int64_t lo = threadIdx.x * per < m ? threadIdx.x * per : m;
  const int64_t hi = lo + per < m ? lo + per : m;
  uint32_t sum = 0;
  for (int64_t e = lo; e < hi; ++e) {
    const int64_t d = e / nt, tt = e % nt;
    sum += cnt[(t0 + tt) * nbuck + d];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(sum, sh, total) + (uint32_t)sstart[sg];
  for (int64_t e = lo; e < hi; ++e) {
    const int64_t d = e / nt, tt = e % nt;
    const int64_t i = (t0 + tt) * nbuck + d;
    const uint32_t c = cnt[i];
    cnt[i] = run;
    run += c;
  }
}

// k_rs_scatter for a segment tile, staged: every 1024-key chunk is first
// ordered by digit in LDS (stable: the wave-multisplit ranks), then written
// out run by run, so consecutive threads write consecutive places of one
// digit's run (coalesced) instead of one scattered element each.  Dynamic
// LDS: 7 nbuck words of counts + the staged keys, values and payload.
constexpr int SEG_CHUNK = 1024;

__global__ void __launch_bounds__(RS_T)
    k_rs_seg_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                     const int64_t* __restrict__ tb, const int64_t* __restrict__ te, int dbits, uint32_t nbuck,
                     const uint32_t* __restrict__ offs, uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                     const uint4* __restrict__ pin, uint4* __restrict__ pout, int pq, const uint32_t* __restrict__ cin,
                     uint32_t* __restrict__ cout) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nb = (int)nbuck;
  uint4* sp = reinterpret_cast<uint4*>(smem_raw);                   // [SEG_CHUNK * pq]
  uint32_t* sk = reinterpret_cast<uint32_t*>(sp + SEG_CHUNK * pq);  // [SEG_CHUNK]
  uint32_t* sv = sk + SEG_CHUNK;                                    // [SEG_CHUNK]
  uint32_t* sc = sv + SEG_CHUNK;                                    // [SEG_CHUNK]
  uint32_t* run = sc + SEG_CHUNK;                                   // [nb] global offsets
  uint32_t* cc = run + nb;                                          // [nb] chunk counts
  uint32_t* lofs = cc + nb;                                         // [nb] chunk digit starts
  uint32_t* wc = lofs + nb;                                         // [4][nb]
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wv = t >> 6;
  for (int d = t; d < nb; d += RS_T) {
    run[d] = offs[(int64_t)tile * nb + d];
    cc[d] = 0;
#pragma unroll
    for (int w = 0; w < RS_T / 64; ++w) wc[w * nb + d] = 0;
  }
  __syncthreads();
  const uint64_t lt = (1ull << lane) - 1ull;
  const int64_t beg = tb[tile], end = te[tile];
  for (int64_t c0 = beg; c0 < end; c0 += SEG_CHUNK) {
    const int clen = (int)(end - c0 < SEG_CHUNK ? end - c0 : SEG_CHUNK);
    // ranks within the chunk: RS_ROUNDS-style rounds of 256 keys
    uint32_t dg[SEG_CHUNK / RS_T], lr[SEG_CHUNK / RS_T];
#pragma unroll
    for (int r = 0; r < SEG_CHUNK / RS_T; ++r) {
      const int i = r * RS_T + t;
      const bool valid = i < clen;
      const uint32_t k = valid ? kin[c0 + i] : 0u;
      const uint32_t d = k % nbuck;
      uint64_t same = __ballot(valid);
      for (int b = 0; b < dbits; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        same &= bit ? m : ~m;
      }
      const int rank = __popcll(same & lt);
      if (valid && rank == 0) wc[wv * nb + d] = (uint32_t)__popcll(same);
      __syncthreads();
      uint32_t pos = 0;
      if (valid) {
        pos = cc[d] + (uint32_t)rank;
        for (int w = 0; w < wv; ++w) pos += wc[w * nb + d];
      }
      __syncthreads();
      if (valid && rank == 0) {
        atomicAdd(&cc[d], wc[wv * nb + d]);
        wc[wv * nb + d] = 0;
      }
      dg[r] = valid ? d : 0xffffffffu;
      lr[r] = pos;
      __syncthreads();
    }
    // digit starts of the chunk (exclusive scan of cc, one thread per
    // digit range)
    {
      const int per = (nb + RS_T - 1) / RS_T;
      const int lo = t * per < nb ? t * per : nb;
      const int hi = lo + per < nb ? lo + per : nb;
      uint32_t ssum = 0;
      for (int d = lo; d < hi; ++d) ssum += cc[d];
      __shared__ uint32_t shs[RS_T];
      uint32_t total;
      uint32_t r0 = block_exclusive_scan(ssum, shs, total);
      for (int d = lo; d < hi; ++d) {
        lofs[d] = r0;
        r0 += cc[d];
      }
    }
    __syncthreads();
    // stage in digit order
#pragma unroll
    for (int r = 0; r < SEG_CHUNK / RS_T; ++r) {
      const int i = r * RS_T + t;
      if (dg[r] != 0xffffffffu) {
        const uint32_t slot = lofs[dg[r]] + lr[r];
        const uint32_t v = vin[c0 + i];
        sk[slot] = kin[c0 + i];
        sv[slot] = v;
        if (pout) {
          for (int q = 0; q < pq; ++q) sp[slot * pq + q] = pin[(int64_t)v * pq + q];
          sc[slot] = cin[v];
        }
      }
    }
    __syncthreads();
    // write runs: slot -> run[digit] + (slot - lofs[digit])
    for (int i = t; i < clen; i += RS_T) {
      const uint32_t k = sk[i];
      const uint32_t d = k % nbuck;
      const uint32_t pos = run[d] + ((uint32_t)i - lofs[d]);
      kout[pos] = k;
      vout[pos] = sv[i];
      if (pout) {
        for (int q = 0; q < pq; ++q) pout[(int64_t)pos * pq + q] = sp[i * pq + q];
        cout[pos] = sc[i];
      }
    }
    __syncthreads();
    for (int d = t; d < nb; d += RS_T) {
      run[d] += cc[d];
      cc[d] = 0;
    }
    __syncthreads();
  }
}

}  // namespace

namespace {
struct SortShape {
  int passes, dbits, nblk, nseg;
  int64_t m;
};

SortShape sort_shape(int64_t n, int key_bits) {
  if (n <= 0 || n >= (1LL << 31)) throw InvalidArgument("radix sort size out of range");
  key_bits = std::max(1, key_bits);
  SortShape h;
  h.passes = (key_bits + 10) / 11;
  h.dbits = (key_bits + h.passes - 1) / h.passes;
  h.nblk = (int)((n + RS_TILE - 1) / RS_TILE);
  h.m = ((int64_t)1 << h.dbits) * h.nblk;
  h.nseg = (int)((h.m + RS_SEG - 1) / RS_SEG);
  return h;
}
}  // namespace

void BinSort::init(int64_t n, int key_bits) {
  const SortShape h = sort_shape(n, key_bits);
  n_ = n;
  passes_ = h.passes;
  dbits_ = h.dbits;
  nblk_ = h.nblk;
  nseg_ = h.nseg;
  cnt_.allocate((size_t)h.m);
  tops_.allocate((size_t)nseg_);
  ktmp_.allocate(n);
  vtmp_.allocate(n);
}

size_t BinSort::workspace_bytes(int64_t n, int key_bits) {
  const SortShape h = sort_shape(n, key_bits);
  size_t b = Carve::rounded((size_t)h.m * 4) + Carve::rounded((size_t)h.nseg * 4);
  if (h.passes > 1) b += 2 * Carve::rounded((size_t)n * 4);
  return b;
}

void BinSort::init(int64_t n, int key_bits, Carve& ws) {
  const SortShape h = sort_shape(n, key_bits);
  n_ = n;
  passes_ = h.passes;
  dbits_ = h.dbits;
  nblk_ = h.nblk;
  nseg_ = h.nseg;
  ws.take(cnt_, (size_t)h.m);
  ws.take(tops_, (size_t)nseg_);
  // single-pass sorts write the outputs directly: no ping-pong buffers
  if (passes_ > 1) {
    ws.take(ktmp_, (size_t)n);
    ws.take(vtmp_, (size_t)n);
  } else {
    ktmp_.release();
    vtmp_.release();
  }
}

void BinSort::set_segments(const std::vector<int64_t>& starts, int nbuck) {
  if (starts.size() < 2 || starts.front() != 0 || starts.back() != n_)
    throw InvalidArgument("sort segments must cover the keys");
  if (nbuck < 1 || nbuck > RS_MAXB) throw InvalidArgument("segment digits must fit 11 bits");
  std::vector<int64_t> tb, te;
  std::vector<int> t0;
  for (size_t i = 0; i + 1 < starts.size(); ++i) {
    if (starts[i + 1] < starts[i]) throw InvalidArgument("sort segments must be ordered");
    t0.push_back((int)tb.size());
    for (int64_t b = starts[i]; b < starts[i + 1]; b += RS_TILE) {
      tb.push_back(b);
      te.push_back(std::min<int64_t>(starts[i + 1], b + RS_TILE));
    }
  }
  t0.push_back((int)tb.size());
  nsegs_ = (int)starts.size() - 1;
  nseg_tiles_ = (int)tb.size();
  seg_buck_ = nbuck;
  tile_beg_.allocate(std::max<size_t>(1, tb.size()));
  tile_end_.allocate(std::max<size_t>(1, te.size()));
  if (!tb.empty()) {
    tile_beg_.upload(tb.data(), tb.size());
    tile_end_.upload(te.data(), te.size());
  }
  seg_tile0_.allocate(t0.size());
  seg_tile0_.upload(t0.data(), t0.size());
  seg_start_.allocate(starts.size());
  seg_start_.upload(starts.data(), starts.size());
  seg_cnt_.allocate(std::max<size_t>(1, (size_t)nseg_tiles_ * nbuck));
  seg_ = true;
}

void BinSort::sort(const uint32_t* keys, const uint32_t* vals, uint32_t* keys_out, uint32_t* vals_out,
                   hipStream_t s, const SortPayload* payload) {
  if (payload && (payload->rec_words <= 0 || payload->rec_words % 4 != 0))
    throw InvalidArgument("sort payload records must be whole 16-byte words");
  const size_t seg_lds = (size_t)SEG_CHUNK * (16 * (payload ? payload->rec_words / 4 : 0) + 12) +
                        7 * (size_t)seg_buck_ * 4;
  if (seg_ && seg_lds <= 64 * 1024) {
    int dbits = 1;
    while ((1 << dbits) < seg_buck_) ++dbits;
    if (nseg_tiles_ > 0) {
      hipLaunchKernelGGL(k_rs_seg_upsweep, dim3(nseg_tiles_), dim3(RS_T), 0, s, keys, tile_beg_.get(),
                         tile_end_.get(), (uint32_t)seg_buck_, seg_cnt_.get());
      hipLaunchKernelGGL(k_rs_seg_scan, dim3(nsegs_), dim3(RS_T), 0, s, seg_cnt_.get(), seg_tile0_.get(),
                         seg_start_.get(), seg_buck_);
      hipLaunchKernelGGL(k_rs_seg_scatter, dim3(nseg_tiles_), dim3(RS_T), seg_lds, s, keys, vals, tile_beg_.get(),
                         tile_end_.get(), dbits, (uint32_t)seg_buck_, seg_cnt_.get(), keys_out, vals_out,
                         payload ? static_cast<const uint4*>(payload->rec_in) : nullptr,
                         payload ? static_cast<uint4*>(payload->rec_out) : nullptr,
                         payload ? payload->rec_words / 4 : 0, payload ? payload->cell_in : nullptr,
                         payload ? payload->cell_out : nullptr);
      FRESCO_HIP_CHECK(hipGetLastError());
    }
    return;
  }
  const uint32_t* kin = keys;
  const uint32_t* vin = vals;
  const int nbuck = 1 << dbits_;
  const uint32_t dmask = (uint32_t)nbuck - 1u;
  const int64_t m = (int64_t)nbuck * nblk_;
  for (int p = 0; p < passes_; ++p) {
    // the last pass writes the outputs; earlier ones alternate so that no
    // pass reads and writes the same buffers
    const bool to_out = ((passes_ - 1 - p) % 2) == 0;
    uint32_t* ko = to_out ? keys_out : ktmp_.get();
    uint32_t* vo = to_out ? vals_out : vtmp_.get();
    const int shift = dbits_ * p;
    hipLaunchKernelGGL(k_rs_upsweep, dim3(nblk_), dim3(RS_T), 0, s, kin, n_, shift, dmask, nbuck, cnt_.get(), nblk_);
    hipLaunchKernelGGL(k_rs_scan_seg, dim3(nseg_), dim3(RS_T), 0, s, cnt_.get(), m, tops_.get());
    hipLaunchKernelGGL(k_rs_scan_top, dim3(1), dim3(RS_T), 0, s, tops_.get(), nseg_);
    const bool last = p == passes_ - 1;
    const SortPayload* pl = last ? payload : nullptr;
    hipLaunchKernelGGL(k_rs_scatter, dim3(nblk_), dim3(RS_T), 0, s, kin, vin, n_, shift, dmask, dbits_, nbuck,
                       cnt_.get(), tops_.get(), nblk_, ko, vo, pl ? static_cast<const uint4*>(pl->rec_in) : nullptr,
                       pl ? static_cast<uint4*>(pl->rec_out) : nullptr, pl ? pl->rec_words / 4 : 0,
                       pl ? pl->cell_in : nullptr, pl ? pl->cell_out : nullptr);
    FRESCO_HIP_CHECK(hipGetLastError());
    kin = ko;
    vin = vo;
  }
}

}  // namespace fresco_hip
