// tm_reduced.hip -- symmetry-reduced transfer matrices on gfx950.
//
// Orbit tables.  The reference enumerates orbits by a sequential scan with a
// visited bitset (TMat/tmatreduce.cpp:19-111, TMat/tmat3d.cpp:17-98): the
// representative of an orbit is its smallest state, representatives come out
// ascending and the duplicate count is the orbit size.  Here every state is
// canonicalised in parallel (minimum over the group images), the "is its own
// minimum" flags are packed into a bitmap, and a prefix sum over the bitmap
// words gives each representative its index -- the same ascending order.
// The orbit size is |G| / |Stab(s)|, counted over the same group images.
//
// Operator.  T_red(o,o') = (d_o d_o')^-1/2 sum_{s in o, s' in o'} T(s,s')
// (the algebra behind XAMMatVecProd2D::perform_op, TMat/tmatreduce.cpp:406-529):
//   y_o = d_o Tvec_o [ (x)K  u ](rep_o),   u(s) = Tvec_{o(s)} x_{o(s)}.
// Generic path: expand u over the full space with a state->orbit map, run
// the Kronecker passes, gather at the representatives.
// Fused path (2D periodic, L = 2n even).  Split every state into its even
// bits E and odd bits O, s = interleave(E, O).  The full transform is
// (x)K over E times (x)K over O, and a rotation of s by two sites rotates E
// and O together, so u(E, O) = u(rotl E, rotl O) and therefore
//   A(E, O) = sum_O' K(O, O') u(E, O')  satisfies  A(rotl E, rotl O) = A(E, O).
// Pass 1 computes A only for columns E that are necklace representatives
// (1,182 of 16,384 at L=28), gathering u from a column-major state->orbit map,
// and stores each column grouped by row necklace (the <= n rotations of a
// row share one 64-byte group).  Pass 2 rebuilds, for every row necklace
// representative O_r, the full E-column A(., O_r) from those groups,
// transforms over E and writes y at one point of every orbit (each orbit has
// an image whose O part is a necklace representative).  The expanded 2^L
// vector never exists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>

#include "tm_device.hpp"

namespace ltm {
namespace {

using dev::pop;

// ---------------------------------------------------------------------------
// group actions
// ---------------------------------------------------------------------------
struct GroupDev {
  int dim, L, s1, s2, bc, censym, rotate, zgroup, ylayer;
  uint64_t mask;
  uint64_t cm1, cmr1;  // 3D column masks for row rotation by 1
  uint64_t rowmask;    // low s2 bits
  int order;           // |G|
};

__device__ __forceinline__ uint64_t row_rot1(uint64_t s, const GroupDev &g)
{
  return ((s >> 1) & g.cmr1) | ((s & g.cm1) << (g.s2 - 1));
}

// reverse the order of the s1 rows (reflectState_3D(.,1), utility.cpp:109-117)
__device__ __forceinline__ uint64_t rows_reverse(uint64_t s, const GroupDev &g)
{
  uint64_t r = 0;
  for (int i = 0; i < g.s1; ++i) {
    r = (r << g.s2) | (s & g.rowmask);
    s >>= g.s2;
  }
  return r;
}
// reverse the bits inside every row (reflectState_3D(.,0), utility.cpp:100-108)
__device__ __forceinline__ uint64_t rows_mirror(uint64_t s, const GroupDev &g)
{
  return rows_reverse(dev::brev(s, g.L), g);
}
// rotateState_3D (utility.cpp:78-91) for a square layer
__device__ __forceinline__ uint64_t layer_rot90(uint64_t s, const GroupDev &g)
{
  const int n = g.s1;
  uint64_t r = 0;
  for (int rp = 0; rp < n; ++rp)
    for (int rq = 0; rq < n; ++rq)
      if (s & (1ull << (rp * n + rq))) r |= 1ull << ((rq + 1) * n - (rp + 1));
  return r;
}

// Visit every group image of s (with multiplicity |G| in total) and return
// the minimum; *stab receives the number of images equal to s.
__device__ uint64_t canon(uint64_t s, const GroupDev &g, int *stab)
{
  uint64_t m = s;
  int st = 0;
  if (g.dim == 2) {
    const int nf = g.censym ? 2 : 1;
    for (int f = 0; f < nf; ++f) {
      const uint64_t s0 = f ? (~s & g.mask) : s;
      if (g.ylayer) {
        // both layers rotated together, and both reflected
        // (TMat/tmatreduce.cpp:165-196)
        const int n = g.ylayer;
        const uint64_t lm = (1ull << n) - 1;
        for (int r = 0; r < 2; ++r) {
          uint64_t t = r ? ((dev::brev(s0 >> n, n) << n) | dev::brev(s0 & lm, n)) : s0;
          for (int k = 0; k < n; ++k) {
            m = t < m ? t : m;
            st += (t == s);
            t = (dev::rotr(t >> n, n, 1) << n) | dev::rotr(t & lm, n, 1);
          }
        }
      } else if (g.zgroup) {
        // translations by two sites t and their mirror images i -> -i,
        // brev(rotr(t, 1)) (TMat/tmatreduce.cpp:272-297)
        uint64_t t = s0;
        for (int k = 0; k < g.L / 2; ++k) {
          const uint64_t r = dev::brev(dev::rotr(t, g.L, 1), g.L);
          m = t < m ? t : m;
          m = r < m ? r : m;
          st += (t == s) + (r == s);
          t = dev::rotr(t, g.L, 2);
        }
      } else if (g.bc == 0) {
        uint64_t t = s0;
        for (int k = 0; k < g.L; ++k) {
          const uint64_t r = dev::brev(t, g.L);
          m = t < m ? t : m;
          m = r < m ? r : m;
          st += (t == s) + (r == s);
          t = dev::rotr(t, g.L, 1);
        }
      } else {
        const uint64_t r = dev::brev(s0, g.L);
        m = s0 < m ? s0 : m;
        m = r < m ? r : m;
        st += (s0 == s) + (r == s);
      }
    }
  } else {
    const int nrot = g.rotate ? 2 : 1, nf = g.censym ? 2 : 1;
    uint64_t sr = s;
    for (int ir = 0; ir < nrot; ++ir) {
      for (int f = 0; f < nf; ++f) {
        const uint64_t sf = f ? (~sr & g.mask) : sr;
        for (int m1 = 0; m1 < 2; ++m1) {
          const uint64_t sm1 = m1 ? rows_mirror(sf, g) : sf;
          for (int m2 = 0; m2 < 2; ++m2) {
            uint64_t t = m2 ? rows_reverse(sm1, g) : sm1;
            for (int a = 0; a < g.s2; ++a) {
              uint64_t u = t;
              for (int b = 0; b < g.s1; ++b) {
                m = u < m ? u : m;
                st += (u == s);
                u = dev::rotr(u, g.L, g.s2);
              }
              t = row_rot1(t, g);
            }
          }
        }
      }
      sr = layer_rot90(sr, g);
    }
  }
  if (stab) *stab = st;
  return m;
}

// bitmap[w] bit j = (state 64w+j is the minimum of its orbit); wcount[w] = popcount
__global__ void k_canon_bits(GroupDev g, uint64_t nstate, uint64_t *bitmap, uint32_t *wcount,
                             uint64_t nwords)
{
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nwords;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t bits = 0;
    const uint64_t s0 = w << 6;
    for (int j = 0; j < 64; ++j) {
      const uint64_t s = s0 + j;
      if (s >= nstate) break;
      if (canon(s, g, nullptr) == s) bits |= 1ull << j;
    }
    bitmap[w] = bits;
    wcount[w] = (uint32_t)pop(bits);
  }
}

// Block-local exclusive scan over 4096-entry chunks of wcount, plus chunk sums.
constexpr int SCAN_CHUNK = 4096;
constexpr int SCAN_THREADS = 256;

__global__ __launch_bounds__(SCAN_THREADS) void k_chunk_sum(const uint32_t *in, uint64_t n,
                                                            uint64_t *sums)
{
  __shared__ uint64_t red[SCAN_THREADS];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
  uint64_t acc = 0;
  for (int i = threadIdx.x; i < SCAN_CHUNK; i += SCAN_THREADS)
    if (base + i < n) acc += in[base + i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = SCAN_THREADS / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(SCAN_THREADS) void k_chunk_scan(const uint32_t *in, uint64_t n,
                                                             const uint64_t *offs,
                                                             uint64_t *out)
{
  // each thread scans 16 consecutive entries; then a block scan of thread sums
  __shared__ uint64_t ts[SCAN_THREADS];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK + threadIdx.x * 16;
  uint64_t loc[16];
  uint64_t acc = 0;
  for (int i = 0; i < 16; ++i) {
    loc[i] = acc;
    if (base + i < n) acc += in[base + i];
  }
  ts[threadIdx.x] = acc;
  __syncthreads();
  // Hillis-Steele inclusive scan of ts
  for (int o = 1; o < SCAN_THREADS; o <<= 1) {
    uint64_t v = threadIdx.x >= o ? ts[threadIdx.x - o] : 0;
    __syncthreads();
    ts[threadIdx.x] += v;
    __syncthreads();
  }
  const uint64_t pre = offs[blockIdx.x] + (threadIdx.x ? ts[threadIdx.x - 1] : 0);
  for (int i = 0; i < 16; ++i)
    if (base + i < n) out[base + i] = pre + loc[i];
}

__global__ void k_write_reps(GroupDev g, const uint64_t *bitmap, const uint64_t *wpre,
                             uint64_t nwords, uint64_t *reps, uint16_t *dups)
{
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nwords;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t bits = bitmap[w];
    uint64_t idx = wpre[w];
    while (bits) {
      const int j = __ffsll((unsigned long long)bits) - 1;
      bits &= bits - 1;
      const uint64_t s = (w << 6) | j;
      int st = 0;
      canon(s, g, &st);
      reps[idx] = s;
      dups[idx] = (uint16_t)(g.order / st);
      ++idx;
    }
  }
}

__device__ __forceinline__ uint32_t rank_of(uint64_t rep, const uint64_t *bitmap,
                                            const uint64_t *wpre)
{
  const uint64_t w = rep >> 6;
  const uint64_t below = bitmap[w] & ((1ull << (rep & 63)) - 1);
  return (uint32_t)(wpre[w] + pop(below));
}

__global__ void k_build_map(GroupDev g, uint64_t nstate, const uint64_t *bitmap,
                            const uint64_t *wpre, uint32_t *map)
{
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < nstate;
       s += (uint64_t)gridDim.x * blockDim.x)
    map[s] = rank_of(canon(s, g, nullptr), bitmap, wpre);
}

// spread the low 16 bits of x to the even bit positions
__device__ __forceinline__ uint64_t spread16(uint32_t x)
{
  uint64_t v = x & 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFull;
  v = (v | (v << 4)) & 0x0F0F0F0Full;
  v = (v | (v << 2)) & 0x33333333ull;
  v = (v | (v << 1)) & 0x55555555ull;
  return v;
}

__global__ void k_build_cmap(GroupDev g, int n, const uint16_t *NR, uint64_t nNk,
                             const uint16_t *order, const uint64_t *bitmap, const uint64_t *wpre,
                             uint32_t *cmap)
{
  const uint64_t total = nNk << n;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t c = i >> n;
    const uint32_t O = order[i & ((1ull << n) - 1)];
    const uint64_t s = spread16(NR[c]) | (spread16(O) << 1);
    cmap[i] = rank_of(canon(s, g, nullptr), bitmap, wpre);
  }
}

GroupDev to_dev(const GroupSpec &gs)
{
  GroupDev g{};
  g.dim = gs.dim;
  g.L = gs.L;
  g.s1 = gs.s1;
  g.s2 = gs.s2;
  g.bc = gs.bc;
  g.censym = gs.censym;
  g.rotate = gs.rotate;
  g.zgroup = gs.zgroup;
  g.ylayer = gs.ylayer;
  g.mask = (1ull << gs.L) - 1;
  if (gs.dim == 3) {
    uint64_t c1 = 0, cr1 = 0;
    for (int r = 0; r < gs.s1; ++r) {
      c1 |= 1ull << (r * gs.s2);
      cr1 |= ((1ull << (gs.s2 - 1)) - 1) << (r * gs.s2);
    }
    g.cm1 = c1;
    g.cmr1 = cr1;
    g.rowmask = (1ull << gs.s2) - 1;
    g.order = gs.s1 * gs.s2 * 4 * (gs.censym ? 2 : 1) * (gs.rotate ? 2 : 1);
  } else if (gs.ylayer) {
    g.order = 2 * gs.ylayer * (gs.censym ? 2 : 1);
  } else if (gs.zgroup) {
    g.order = gs.L * (gs.censym ? 2 : 1);
  } else {
    g.order = (gs.bc == 0 ? 2 * gs.L : 2) * (gs.censym ? 2 : 1);
  }
  return g;
}

unsigned grid_for(uint64_t n, int threads)
{
  uint64_t b = (n + threads - 1) / threads;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Exclusive prefix of the word counts -> wpre (device), returns the total.
uint64_t scan_words(const uint32_t *wcount, uint64_t nwords, uint64_t *wpre, hipStream_t s)
{
  const uint64_t nchunk = (nwords + SCAN_CHUNK - 1) / SCAN_CHUNK;
  DevBuf sums(nchunk * sizeof(uint64_t)), offs(nchunk * sizeof(uint64_t));
  hipLaunchKernelGGL(k_chunk_sum, dim3((unsigned)nchunk), dim3(SCAN_THREADS), 0, s, wcount,
                     nwords, sums.as<uint64_t>());
  TM_LAUNCH_CHECK();
  std::vector<uint64_t> h(nchunk), o(nchunk);
  TM_HIP(hipMemcpyAsync(h.data(), sums.get(), nchunk * 8, hipMemcpyDeviceToHost, s));
  TM_HIP(hipStreamSynchronize(s));
  uint64_t acc = 0;
  for (uint64_t i = 0; i < nchunk; ++i) { o[i] = acc; acc += h[i]; }
  TM_HIP(hipMemcpyAsync(offs.get(), o.data(), nchunk * 8, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_chunk_scan, dim3((unsigned)nchunk), dim3(SCAN_THREADS), 0, s, wcount,
                     nwords, offs.as<uint64_t>(), wpre);
  TM_LAUNCH_CHECK();
  TM_HIP(hipStreamSynchronize(s));
  return acc;
}

struct OrbitScratch {
  DevBuf bitmap, wcount, wpre;
  uint64_t nwords = 0;
};

uint64_t orbit_bitmap(const GroupDev &g, OrbitScratch &sc, hipStream_t s)
{
  const uint64_t nstate = 1ull << g.L;
  sc.nwords = (nstate + 63) / 64;
  sc.bitmap.alloc(sc.nwords * 8);
  sc.wcount.alloc(sc.nwords * 4);
  sc.wpre.alloc(sc.nwords * 8);
  hipLaunchKernelGGL(k_canon_bits, dim3(grid_for(sc.nwords, 256)), dim3(256), 0, s, g, nstate,
                     sc.bitmap.as<uint64_t>(), sc.wcount.as<uint32_t>(), sc.nwords);
  TM_LAUNCH_CHECK();
  return scan_words(sc.wcount.as<uint32_t>(), sc.nwords, sc.wpre.as<uint64_t>(), s);
}

}  // namespace

namespace {
// Expanded-space vector of a reduced one: xf[s] = xr[o(s)] / sqrt(|o(s)|),
// with o(s) found by canonicalising s and a binary search of the ascending
// representatives.  Inner products and the transfer matrix carry over:
// <xf, T yf> = <xr, T_red yr>.
__global__ void k_expand_orbits(GroupDev g, uint64_t nstate, const uint64_t *__restrict__ reps,
                                const uint16_t *__restrict__ dups, int64_t M,
                                const double *__restrict__ xr, double *__restrict__ xf)
{
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < nstate;
       s += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t c = canon(s, g, nullptr);
    int64_t lo = 0, hi = M - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (reps[mid] < c) lo = mid + 1;
      else hi = mid;
    }
    xf[s] = xr[lo] * rsqrt((double)dups[lo]);
  }
}
}  // namespace

void expand_orbits(const GroupSpec &gs, const uint64_t *reps, const uint16_t *dups, int64_t M,
                   const double *xr, double *xf, hipStream_t s)
{
  const GroupDev g = to_dev(gs);
  const uint64_t nstate = 1ull << gs.L;
  hipLaunchKernelGGL(k_expand_orbits, dim3(grid_for(nstate, 256)), dim3(256), 0, s, g, nstate,
                     reps, dups, M, xr, xf);
  TM_LAUNCH_CHECK();
}

int64_t enumerate_orbits(const GroupSpec &gs, DevBuf &reps, DevBuf &dups, DevBuf *map,
                         hipStream_t s)
{
  TM_CHECK(gs.L >= 4 && gs.L <= 32, TM_EINVAL, "orbit enumeration supports 4 <= L <= 32");
  const GroupDev g = to_dev(gs);
  OrbitScratch sc;
  const uint64_t M = orbit_bitmap(g, sc, s);
  TM_CHECK(M > 0 && M < (1ull << 32), TM_EINVAL, "orbit count out of range");
  reps.alloc(M * sizeof(uint64_t));
  dups.alloc(M * sizeof(uint16_t));
  hipLaunchKernelGGL(k_write_reps, dim3(grid_for(sc.nwords, 256)), dim3(256), 0, s, g,
                     sc.bitmap.as<uint64_t>(), sc.wpre.as<uint64_t>(), sc.nwords,
                     reps.as<uint64_t>(), dups.as<uint16_t>());
  TM_LAUNCH_CHECK();
  if (map) {
    const uint64_t nstate = 1ull << gs.L;
    map->alloc(nstate * sizeof(uint32_t));
    hipLaunchKernelGGL(k_build_map, dim3(grid_for(nstate, 256)), dim3(256), 0, s, g, nstate,
                       sc.bitmap.as<uint64_t>(), sc.wpre.as<uint64_t>(), map->as<uint32_t>());
    TM_LAUNCH_CHECK();
  }
  TM_HIP(hipStreamSynchronize(s));
  return (int64_t)M;
}

namespace {
__global__ void k_remap_u32(uint32_t *__restrict__ a, const uint32_t *__restrict__ q, uint64_t n)
{
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    a[i] = q[a[i]];
}
}  // namespace

// LDS position of element e of a fused-pass column of n bits: the common
// swizzle, plus (at n = 14) bits 12-13 folded into bank bits 3-4, which
// keeps both layouts of the two-round transform (fused_kron14) free of bank
// conflicts.  XOR-linear and a bijection on [0, 2^n).
__host__ __device__ __forceinline__ int fsw(int e, int n)
{
  return dev::sw(e) ^ (n == 14 ? ((e >> 12) & 3) << 3 : 0);
}

void build_fused_tables(const GroupSpec &gs, const DevBuf &reps, int64_t M, FusedTables &ft,
                        hipStream_t s)
{
  TM_CHECK(gs.dim == 2 && gs.bc == 0 && (gs.L % 2) == 0 && gs.L >= 8 && gs.L <= 30, TM_EINVAL,
           "fused reduced path needs a periodic 2D layer of even width 8..30");
  const int n = gs.L / 2;
  const uint32_t nw = 1u << n, wmask = nw - 1;
  ft.n = n;
  auto rotl = [&](uint32_t w, int d) -> uint32_t {
    d %= n;
    return d ? (((w << d) & wmask) | (w >> (n - d))) : w;
  };
  // necklaces of n-bit words
  std::vector<uint16_t> nk(nw), NR;
  std::vector<uint8_t> rot(nw), per;
  std::vector<int32_t> repidx(nw, -1);
  for (uint32_t w = 0; w < nw; ++w) {
    uint32_t m = w;
    for (int j = 1; j < n; ++j) m = std::min(m, rotl(w, j));
    if (m == w) {
      repidx[w] = (int32_t)NR.size();
      NR.push_back((uint16_t)w);
      int p = 1;
      while (rotl(w, p) != w) ++p;
      per.push_back((uint8_t)p);
    }
  }
  for (uint32_t w = 0; w < nw; ++w) {
    uint32_t m = w;
    for (int j = 1; j < n; ++j) m = std::min(m, rotl(w, j));
    const int r = repidx[m];
    nk[w] = (uint16_t)r;
    int k = 0;
    while (rotl(m, k) != w) ++k;
    rot[w] = (uint8_t)k;
  }
  ft.nNk = (int)NR.size();
  const int nNk = ft.nNk;
  // the pass kernels cover a column (2^n entries, 4 per group) and a row
  // (4 quarters per class) in one round of their thread's registers
  TM_CHECK(n <= 14 && nNk * 4 <= 5 * 1024, TM_EINVAL, "fused reduced path needs L <= 28");
  std::vector<uint16_t> revw(nw);
  for (uint32_t w = 0; w < nw; ++w) {
    uint32_t r = 0;
    for (int i = 0; i < n; ++i) r |= ((w >> i) & 1u) << (n - 1 - i);
    revw[w] = (uint16_t)r;
  }
  // mirrors: the site mirror i -> -i maps the E word to rotl_1(rev E) and
  // the O word to rev O; both land in the necklace class of the reversed word
  std::vector<uint16_t> mcl(nNk);
  for (int c = 0; c < nNk; ++c) mcl[c] = nk[revw[NR[c]]];
  // stored columns and rows: the bracelet representatives (c <= its mirror)
  std::vector<uint16_t> colS, rowS;
  std::vector<int32_t> sidx(nNk, -1);
  for (int c = 0; c < nNk; ++c)
    if (c <= mcl[c]) {
      sidx[c] = (int32_t)colS.size();
      colS.push_back((uint16_t)c);
      rowS.push_back((uint16_t)c);
    }
  ft.nCol = (int)colS.size();
  ft.nRow = (int)rowS.size();
  std::vector<uint16_t> csrc(nNk);
  std::vector<uint8_t> crot(nNk, 255), rrot(nNk, 0);
  for (int c = 0; c < nNk; ++c) {
    if (sidx[c] >= 0) {
      csrc[c] = (uint16_t)sidx[c];
    } else {
      const int src = mcl[c];
      csrc[c] = (uint16_t)sidx[src];
      const uint32_t w = rotl(revw[NR[src]], 1);
      int m = 0;
      while (m < per[c] && rotl(w, m) != NR[c]) ++m;
      TM_CHECK(m < per[c], TM_EINVAL, "mirror column has no rotation onto its representative");
      crot[c] = (uint8_t)m;
    }
    const uint32_t rw = revw[NR[c]];
    int t = 0;
    while (t < per[c] && rotl(NR[mcl[c]], t) != rw) ++t;
    TM_CHECK(t < per[c], TM_EINVAL, "mirror row has no rotation onto the reversed word");
    rrot[c] = (uint8_t)t;
  }
  // column entries: O words necklace-major (class, then rotation), so that
  // the entries of one O class are consecutive
  std::vector<uint16_t> order;
  std::vector<uint16_t> qof;
  order.reserve(nw);
  qof.reserve(nw);
  for (int q = 0; q < nNk; ++q)
    for (int k = 0; k < per[q]; ++k) {
      order.push_back((uint16_t)rotl(NR[q], k));
      qof.push_back((uint16_t)q);
    }
  // slot k of row r holds the O word that pass 2 needs for the rotation
  // E = rotl_k(NR[c]) of a stored column: A(NR[c], rotr_k NR[r]) = A(NR[c],
  // rotl_{(n-k) mod p_r} NR[r]), k < n (k >= p_c is never read)
  std::vector<uint16_t> slot((size_t)nNk * 16, 0xFFFF);
  for (int r = 0; r < nNk; ++r)
    for (int k = 0; k < n; ++k)
      slot[(size_t)r * 16 + k] = (uint16_t)fsw(rotl(NR[r], (n - k) % per[r]), n);
  // evaluation point of every orbit (see FusedTables): the images of the
  // representative under the one-site shift, the mirror and the flip carry
  // every (E class, O class) pair of the orbit
  std::vector<uint64_t> hreps(M);
  copy_sync(hreps.data(), reps.get(), M * 8, hipMemcpyDeviceToHost, s);
  std::vector<int32_t> rowq(nNk, -1);
  for (int q = 0; q < ft.nRow; ++q) rowq[rowS[q]] = q;
  auto compact = [&](uint64_t v) {  // bits 0,2,4,... -> 0,1,2,...
    uint32_t r = 0;
    for (int i = 0; i < n; ++i) r |= (uint32_t)((v >> (2 * i)) & 1) << i;
    return r;
  };
  const bool flip = gs.censym != 0;
  std::vector<int64_t> cnt(ft.nRow + 1, 0);
  std::vector<uint16_t> pr(M), pe(M), pkey(M);
  for (int64_t o = 0; o < M; ++o) {
    uint32_t iE[8], iO[8];
    int ni = 0;
    const uint32_t E0 = compact(hreps[o]), O0 = compact(hreps[o] >> 1);
    for (int sh = 0; sh < 2; ++sh) {
      // one-site shift: (E, O) -> (rotl_1 O, E)
      const uint32_t E1 = sh ? rotl(O0, 1) : E0, O1 = sh ? E0 : O0;
      for (int mi = 0; mi < 2; ++mi) {
        const uint32_t E2 = mi ? rotl(revw[E1], 1) : E1, O2 = mi ? revw[O1] : O1;
        for (int f = 0; f < (flip ? 2 : 1); ++f) {
          iE[ni] = f ? (~E2 & wmask) : E2;
          iO[ni] = f ? (~O2 & wmask) : O2;
          ++ni;
        }
      }
    }
    // stored O classes of the orbit, ascending and distinct
    int B[8], nb = 0;
    for (int i = 0; i < ni; ++i) {
      const int b = nk[iO[i]];
      if (rowq[b] < 0) continue;
      bool seen = false;
      for (int j = 0; j < nb; ++j) seen |= B[j] == b;
      if (!seen) B[nb++] = b;
    }
    TM_CHECK(nb > 0, TM_EINVAL, "orbit has no evaluation point in a stored row");
    std::sort(B, B + nb);
    int sum = 0;
    for (int j = 0; j < nb; ++j) sum += B[j];
    const int R = B[sum % nb];
    int key = nNk, pick = -1;
    for (int i = 0; i < ni; ++i)
      if (nk[iO[i]] == R) {
        key = std::min(key, (int)nk[iE[i]]);
        if (pick < 0) pick = i;
      }
    const int k = rot[iO[pick]];
    pr[o] = (uint16_t)rowq[R];
    pe[o] = (uint16_t)rotl(iE[pick], n - k);  // rotate by k two-site steps back
    pkey[o] = (uint16_t)key;
    ++cnt[rowq[R] + 1];
  }
  for (int q = 0; q < ft.nRow; ++q) cnt[q + 1] += cnt[q];
  std::vector<uint32_t> olist(M);
  std::vector<uint16_t> elist(M);
  std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
  for (int64_t o = 0; o < M; ++o) olist[pos[pr[o]]++] = (uint32_t)o;
  for (int q = 0; q < ft.nRow; ++q)
    std::sort(olist.begin() + cnt[q], olist.begin() + cnt[q + 1], [&](uint32_t a, uint32_t b) {
      if (pkey[a] != pkey[b]) return pkey[a] < pkey[b];
      const uint32_t ea = pe[a], eb = pe[b];
      if (nk[ea] != nk[eb]) return nk[ea] < nk[eb];
      if (rot[ea] != rot[eb]) return rot[ea] < rot[eb];
      return a < b;
    });
  for (int64_t i = 0; i < M; ++i) elist[i] = (uint16_t)fsw(pe[olist[i]], n);
  // pass-2 rows, the ones with most orbits launched first
  std::vector<int> rorder(ft.nRow);
  for (int q = 0; q < ft.nRow; ++q) rorder[q] = q;
  std::stable_sort(rorder.begin(), rorder.end(), [&](int a, int b) {
    return cnt[a + 1] - cnt[a] > cnt[b + 1] - cnt[b];
  });
  std::vector<uint64_t> rinfo(ft.nRow);
  for (int i = 0; i < ft.nRow; ++i) {
    const int q = rorder[i], r = rowS[q];
    rinfo[i] = (uint64_t)r | ((uint64_t)mcl[r] << 16) | ((uint64_t)rrot[r] << 32) |
               ((uint64_t)q << 40);
  }
  // pass-2 items: stored columns, then the mirrored classes by source column
  std::vector<int> mirc;
  for (int c = 0; c < nNk; ++c)
    if (crot[c] != 255) mirc.push_back(c);
  std::stable_sort(mirc.begin(), mirc.end(), [&](int a, int b) { return csrc[a] < csrc[b]; });
  std::vector<uint64_t> itab;
  itab.reserve(nNk);
  for (int i = 0; i < ft.nCol; ++i) {
    const int c = colS[i];
    itab.push_back((uint64_t)i | ((uint64_t)NR[c] << 16) | ((uint64_t)per[c] << 32) |
                   (255ull << 40));
  }
  for (int c : mirc)
    itab.push_back((uint64_t)csrc[c] | ((uint64_t)NR[c] << 16) | ((uint64_t)per[c] << 32) |
                   ((uint64_t)crot[c] << 40));
  TM_CHECK((int)itab.size() == nNk, TM_EINVAL, "pass-2 item table does not cover every class");
  auto up = [&](DevBuf &d, const void *src, size_t bytes) {
    d.alloc(bytes);
    copy_sync(d.get(), src, bytes, hipMemcpyHostToDevice, s);
  };
  up(ft.slotO, slot.data(), slot.size() * 2);
  up(ft.ostart, cnt.data(), cnt.size() * 8);
  up(ft.olist, olist.data(), olist.size() * 4);
  up(ft.Elist, elist.data(), elist.size() * 2);
  up(ft.itab, itab.data(), itab.size() * 8);
  up(ft.rinfo, rinfo.data(), rinfo.size() * 8);
  // internal orbit numbering = position in olist (rows contiguous)
  std::vector<uint32_t> inv(M);
  for (int64_t i = 0; i < M; ++i) inv[olist[i]] = (uint32_t)i;
  DevBuf invd, orderd;
  up(invd, inv.data(), inv.size() * 4);
  up(orderd, order.data(), order.size() * 2);
  ft.perm = std::move(olist);

  // column map over the stored necklace columns, in internal numbering
  const GroupDev g = to_dev(gs);
  OrbitScratch sc;
  const uint64_t M2 = orbit_bitmap(g, sc, s);
  TM_CHECK((int64_t)M2 == M, TM_EINVAL, "orbit count mismatch while building fused tables");
  std::vector<uint16_t> colw(ft.nCol);
  for (int i = 0; i < ft.nCol; ++i) colw[i] = NR[colS[i]];
  DevBuf colwd, cmapd;
  up(colwd, colw.data(), colw.size() * 2);
  const uint64_t total = (uint64_t)ft.nCol << n;
  cmapd.alloc(total * 4);
  hipLaunchKernelGGL(k_build_cmap, dim3(grid_for(total, 256)), dim3(256), 0, s, g, n,
                     colwd.as<uint16_t>(), (uint64_t)ft.nCol, orderd.as<uint16_t>(),
                     sc.bitmap.as<uint64_t>(), sc.wpre.as<uint64_t>(), cmapd.as<uint32_t>());
  TM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_remap_u32, dim3(grid_for(total, 256)), dim3(256), 0, s,
                     cmapd.as<uint32_t>(), invd.as<uint32_t>(), total);
  TM_LAUNCH_CHECK();
  std::vector<uint32_t> cmap(total);
  copy_sync(cmap.data(), cmapd.get(), total * 4, hipMemcpyDeviceToHost, s);
  // Stored position of entry j: a wave's 64 threads load 4 consecutive
  // positions each (one group), and gather instruction e of the group
  // should cover 64 consecutive entries (a few runs of x~, not one entry in
  // four of 256): in each block of B = min(256, 2^n) entries, entry
  // e (B/4) + l sits at position 4 l + e.
  const uint32_t Bq = std::min<uint32_t>(256, nw) / 4;
  auto spos = [&](uint32_t j) {
    const uint32_t blk = j / (4 * Bq), rem = j % (4 * Bq);
    return blk * 4 * Bq + 4 * (rem % Bq) + rem / Bq;
  };
  // compressed: base per (column, O class), 8-bit offset per entry
  std::vector<uint32_t> cbase((size_t)ft.nCol * nNk, 0xFFFFFFFFu);
  std::vector<uint8_t> coff(total);
  for (int c = 0; c < ft.nCol; ++c) {
    const uint32_t *mrow = cmap.data() + ((size_t)c << n);
    uint32_t *b = cbase.data() + (size_t)c * nNk;
    for (uint32_t j = 0; j < nw; ++j) b[qof[j]] = std::min(b[qof[j]], mrow[j]);
    for (uint32_t j = 0; j < nw; ++j) {
      const uint32_t d = mrow[j] - b[qof[j]];
      TM_CHECK(d < 256, TM_EINVAL, "column map run longer than 256 orbits");
      coff[((size_t)c << n) + spos(j)] = (uint8_t)d;
    }
  }
  std::vector<uint32_t> jtab(nw);
  for (uint32_t j = 0; j < nw; ++j)
    jtab[spos(j)] = (uint32_t)fsw(order[j], n) | ((uint32_t)qof[j] << 16);
  up(ft.cbase, cbase.data(), cbase.size() * 4);
  up(ft.coff, coff.data(), coff.size());
  up(ft.jtab, jtab.data(), jtab.size() * 4);
}

// ---------------------------------------------------------------------------
// reduced operator application
// ---------------------------------------------------------------------------
namespace {

// xt[o] = (C)(x[o] * scale) * tvec[o]
template <typename C, typename TI>
__global__ void k_xtilde(const TI *__restrict__ x, const C *__restrict__ tvec, C *__restrict__ xt,
                         int64_t M, const double *__restrict__ scale)
{
  const C sc = scale ? (C)(*scale) : (C)1;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < M;
       o += (int64_t)gridDim.x * blockDim.x)
    xt[o] = (C)x[o] * sc * tvec[o];
}

template <typename C>
__global__ void k_expand(const uint32_t *__restrict__ map, const C *__restrict__ xt,
                         C *__restrict__ u, uint64_t nstate)
{
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < nstate;
       s += (uint64_t)gridDim.x * blockDim.x)
    u[s] = xt[map[s]];
}

// expansion times the two-layer weight of the state (c = high layer)
template <typename C>
__global__ void k_expand_w(const uint32_t *__restrict__ map, const C *__restrict__ xt,
                           C *__restrict__ u, uint64_t nstate, int Ly, const C *__restrict__ wtab,
                           int ybm)
{
  __shared__ C stab[320];
  for (int i = threadIdx.x; i < (ybm ? 320 : 192); i += blockDim.x) stab[i] = wtab[i];
  __syncthreads();
  const uint32_t lm = (1u << Ly) - 1;
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < nstate;
       s += (uint64_t)gridDim.x * blockDim.x)
    u[s] = xt[map[s]] * dev::w_layers<C>((uint32_t)(s >> Ly), (uint32_t)s & lm, Ly, stab, ybm);
}

// gather at the representatives, times their two-layer weight
template <typename C, typename TO>
__global__ void k_gather_w(const C *__restrict__ u, const uint64_t *__restrict__ reps,
                           const C *__restrict__ dt, TO *__restrict__ y, int64_t M,
                           const double *__restrict__ scale, int Ly, const C *__restrict__ wtab,
                           int ybm)
{
  __shared__ C stab[320];
  for (int i = threadIdx.x; i < (ybm ? 320 : 192); i += blockDim.x) stab[i] = wtab[i];
  __syncthreads();
  const C inv = scale ? (C)(1.0 / *scale) : (C)1;
  const uint32_t lm = (1u << Ly) - 1;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < M;
       o += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t r = reps[o];
    y[o] = (TO)(dt[o] * u[r] * inv *
                dev::w_layers<C>((uint32_t)(r >> Ly), (uint32_t)r & lm, Ly, stab, ybm));
  }
}

template <typename C, typename TO>
__global__ void k_gather(const C *__restrict__ u, const uint64_t *__restrict__ reps,
                         const C *__restrict__ dt, TO *__restrict__ y, int64_t M,
                         const double *__restrict__ scale)
{
  const C inv = scale ? (C)(1.0 / *scale) : (C)1;
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < M;
       o += (int64_t)gridDim.x * blockDim.x)
    y[o] = (TO)(dt[o] * u[reps[o]] * inv);
}

// Plain in-LDS Kronecker transform over the full space for small L (generic
// path): one workgroup per vector; K = a [[1, r], [r, 1]], f = a^L.
template <typename C>
__global__ __launch_bounds__(1024) void k_kron_small(C *u, int L, C r, C f)
{
  extern __shared__ __align__(16) unsigned char smem[];
  C *lds = reinterpret_cast<C *>(smem);
  const int n = 1 << L;
  for (int i = threadIdx.x; i < n; i += blockDim.x) lds[dev::sw(i)] = u[i];
  __syncthreads();
  dev::lds_kron(lds, L, r);
  for (int i = threadIdx.x; i < n; i += blockDim.x) u[i] = lds[dev::sw(i)] * f;
}

constexpr int FUSED_THREADS = 1024;
// waves per SIMD the pass kernels are compiled for: two 1024-thread blocks
// per CU with FP32 columns (64 KB of LDS each), one with FP64 (128 KB)
template <typename C>
constexpr int fused_waves()
{
  return sizeof(C) == 4 ? 8 : 4;
}

__device__ __forceinline__ void ld4(const float *p, float *v)
{
  const float4 a = *reinterpret_cast<const float4 *>(p);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
  v[3] = a.w;
}
__device__ __forceinline__ void ld4(const double *p, double *v)
{
  const double2 a = reinterpret_cast<const double2 *>(p)[0];
  const double2 b = reinterpret_cast<const double2 *>(p)[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}
__device__ __forceinline__ void st4(float *p, const float *v)
{
  *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void st4(double *p, const double *v)
{
  reinterpret_cast<double2 *>(p)[0] = make_double2(v[0], v[1]);
  reinterpret_cast<double2 *>(p)[1] = make_double2(v[2], v[3]);
}

// k_xtilde over 4-element packs (x, tvec and xt 16-byte aligned); the M mod 4
// tail by the first threads of block 0
template <typename C, typename TI>
__global__ void k_xtilde4(const TI *__restrict__ x, const C *__restrict__ tvec, C *__restrict__ xt,
                          int64_t M, const double *__restrict__ scale)
{
  const C sc = scale ? (C)(*scale) : (C)1;
  const int64_t M4 = M / 4;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < M4;
       p += (int64_t)gridDim.x * blockDim.x) {
    TI a[4];
    C t[4], o[4];
    ld4(x + 4 * p, a);
    ld4(tvec + 4 * p, t);
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = (C)a[i] * sc * t[i];
    st4(xt + 4 * p, o);
  }
  if (blockIdx.x == 0 && threadIdx.x < M - 4 * M4) {
    const int64_t i = 4 * M4 + threadIdx.x;
    xt[i] = (C)x[i] * sc * tvec[i];
  }
}

// Lane-exchange of a value across lane bit B (0..3) inside a 16-lane DPP
// row: quad permutations for bits 0 and 1; row shifts by 4 / 8 for bits 2
// and 3, each written only to the lanes whose partner it reaches (bank
// masks: lanes 0-3, 8-11 of a row take lane + 4, lanes 4-7, 12-15 take
// lane - 4; lanes 0-7 take lane + 8, lanes 8-15 take lane - 8).
template <int CTRL, int BM>
__device__ __forceinline__ float dpp_upd(float old, float x)
{
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(x), CTRL, 0xF, BM, false));
}
template <int CTRL, int BM>
__device__ __forceinline__ double dpp_upd(double old, double x)
{
  const long long o = __double_as_longlong(old), v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)v, CTRL, 0xF, BM, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(v >> 32), CTRL, 0xF, BM, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int B, typename C>
__device__ __forceinline__ C lane_partner(C x)
{
  if constexpr (B == 0) {
    return dpp_upd<0xB1, 0xF>(x, x);  // quad_perm [1, 0, 3, 2]
  } else if constexpr (B == 1) {
    return dpp_upd<0x4E, 0xF>(x, x);  // quad_perm [2, 3, 0, 1]
  } else if constexpr (B == 2) {
    const C p = dpp_upd<0x104, 0x5>(x, x);  // row_shl:4 into banks 0, 2
    return dpp_upd<0x114, 0xA>(p, x);       // row_shr:4 into banks 1, 3
  } else {
    const C p = dpp_upd<0x108, 0x3>(x, x);  // row_shl:8 into banks 0, 1
    return dpp_upd<0x118, 0xC>(p, x);       // row_shr:8 into banks 2, 3
  }
}

// K' = [[1, r], [r, 1]] on all 14 bits of an LDS column (element e at
// lds[fsw(e, 14)]), 1024 threads x 16 values, in two rounds instead of four:
//   A: registers = bits 0-3, lane bits 0-3 = bits 4-7 (DPP), the other lane
//      and wave bits = bits 8-13;
//   B: registers = bits 8-11, lane bits 0-1 = bits 12-13 (DPP), the other
//      lane and wave bits = bits 0-7.
// Each thread writes back the positions it read, so a round needs no
// barrier between its read and its write.  Ends with a block barrier.
template <typename C>
__device__ __forceinline__ void fused_kron14(C *lds, C r)
{
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  C v[16];
  {
    const int sa = fsw(((l & 15) << 4) | ((l >> 4) << 8) | (w << 10), 14);
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = lds[sa ^ i];
    dev::kron_regs<4>(v, r);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      C x = v[i];
      x = x + r * lane_partner<0>(x);
      x = x + r * lane_partner<1>(x);
      x = x + r * lane_partner<2>(x);
      x = x + r * lane_partner<3>(x);
      v[i] = x;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) lds[sa ^ i] = v[i];
  }
  __syncthreads();
  {
    const int sb = fsw(((l & 3) << 12) | (l >> 2) | (w << 4), 14);
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = lds[sb ^ fsw(i << 8, 14)];
    dev::kron_regs<4>(v, r);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      C x = v[i];
      x = x + r * lane_partner<0>(x);
      x = x + r * lane_partner<1>(x);
      v[i] = x;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) lds[sb ^ fsw(i << 8, 14)] = v[i];
  }
  __syncthreads();
}

// the column transform of the fused passes: two rounds at n = 14 (1024
// threads), the four-bit LDS rounds of lds_kron otherwise
template <typename C>
__device__ __forceinline__ void fused_kron(C *lds, int n, C r)
{
  if (n == 14) fused_kron14(lds, r);
  else dev::lds_kron(lds, n, r);
}

// Pass 1: one stored column (E = the necklace representative of a bracelet);
// u(E, O') = xt[cbase[c][q] + coff[c][j]] for every entry j (O' of class q,
// LDS index in jtab); transform over O; store A(E, O) row-major, 16 slots
// per (row necklace, column).  The column's bases are staged in LDS after
// the 2^n values.
template <typename C>
__global__ __launch_bounds__(FUSED_THREADS, fused_waves<C>()) void k_fused_pass1(
    const uint32_t *__restrict__ jtab, const uint32_t *__restrict__ cbase,
    const uint8_t *__restrict__ coff, const C *__restrict__ xt, C *__restrict__ Ap,
    const uint16_t *__restrict__ slotO, int n, int nCol, int nNk, C kr, C kf)
{
  extern __shared__ __align__(16) unsigned char smem[];
  C *lds = reinterpret_cast<C *>(smem);
  uint32_t *lbase = reinterpret_cast<uint32_t *>(smem + (sizeof(C) << n));
  // neighbouring columns write neighbouring 16-slot groups of every row of
  // A and gather neighbouring runs of x~: keep them on one XCD (one L2)
  const int c = (int)dev::xcd_remap(blockIdx.x, nCol);
  if (c < 0) return;
  // groups of 4 consecutive entries: one 16-byte jtab load and one 4-byte
  // offset load each, issued before the bases are staged; one round of
  // UNG groups per thread covers the column (2^n <= 4 UNG FUSED_THREADS,
  // n <= 14 on the fused path)
  const int ng = 1 << (n - 2);
  const uint4 *jt = reinterpret_cast<const uint4 *>(jtab);
  const uint32_t *co = reinterpret_cast<const uint32_t *>(coff + ((size_t)c << n));
  constexpr int UNG = 4;
  const int g0 = threadIdx.x;
  uint4 jj[UNG];
  uint32_t oo[UNG];
#pragma unroll
  for (int u = 0; u < UNG; ++u) {
    const int g = g0 + u * FUSED_THREADS;
    jj[u] = g < ng ? jt[g] : make_uint4(0, 0, 0, 0);
    oo[u] = g < ng ? co[g] : 0u;
  }
  const uint32_t *cb = cbase + (size_t)c * nNk;
  for (int i = threadIdx.x; i < nNk; i += FUSED_THREADS) lbase[i] = cb[i];
  __syncthreads();
  uint32_t idx[4 * UNG];
#pragma unroll
  for (int u = 0; u < UNG; ++u) {
    idx[4 * u + 0] = lbase[jj[u].x >> 16] + (oo[u] & 255u);
    idx[4 * u + 1] = lbase[jj[u].y >> 16] + ((oo[u] >> 8) & 255u);
    idx[4 * u + 2] = lbase[jj[u].z >> 16] + ((oo[u] >> 16) & 255u);
    idx[4 * u + 3] = lbase[jj[u].w >> 16] + (oo[u] >> 24);
  }
  C v[4 * UNG];
#pragma unroll
  for (int i = 0; i < 4 * UNG; ++i) v[i] = xt[idx[i]];
  // the output phase's slot-table words, loaded while the gathers are in
  // flight (one round of UNR items per thread covers nNk*4 <= 5 FUSED_THREADS)
  const int nt = nNk * 4;
  const uint2 *sl = reinterpret_cast<const uint2 *>(slotO);
  constexpr int UNR = 5;
  uint2 w[UNR];
#pragma unroll
  for (int u = 0; u < UNR; ++u) {
    const int t = threadIdx.x + u * FUSED_THREADS;
    w[u] = t < nt ? sl[t] : make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
  }
#pragma unroll
  for (int u = 0; u < UNG; ++u) {
    if (g0 + u * FUSED_THREADS < ng) {
      lds[jj[u].x & 0xFFFFu] = v[4 * u + 0];
      lds[jj[u].y & 0xFFFFu] = v[4 * u + 1];
      lds[jj[u].z & 0xFFFFu] = v[4 * u + 2];
      lds[jj[u].w & 0xFFFFu] = v[4 * u + 3];
    }
  }
  __syncthreads();
  fused_kron(lds, n, kr);
  // output (scaled by kf = a^n): item t = (row necklace r, slot quarter),
  // four slots from one 8-byte slot-table word, one vector store each
#pragma unroll
  for (int u = 0; u < UNR; ++u) {
    const int t = threadIdx.x + u * FUSED_THREADS;
    if (t >= nt) continue;
    const uint32_t o[4] = {w[u].x & 0xFFFFu, w[u].x >> 16, w[u].y & 0xFFFFu, w[u].y >> 16};
    C vv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) vv[i] = o[i] != 0xFFFFu ? lds[o[i]] * kf : (C)0;
    st4(Ap + ((size_t)(t >> 2) * nCol + c) * 16 + 4 * (t & 3), vv);
  }
}

// Pass 2: one stored row necklace O_r = NR[r]; rebuild A(E, O_r) for every E
// and transform over E; write y at the orbits whose evaluation point lies on
// this row.  Item i of itab is a necklace class c with its source column:
// for a stored class, E = rotl_k NR[c] is slot k of A[r][c's column]; for a
// mirrored class (NR[c] = rotl_m(rotl_1(rev NR[src]))) the mirror and
// rotation invariance of A give
//   A(rotr_{m+t+j} NR[c], O_r) = slot j of A[mirror row of r][src],
// with rev(NR[r]) = rotl_t NR[mirror row].  Row r of A and the part of the
// mirror row that the mirrored classes use are read as contiguous runs.
template <typename C, typename TO>
__global__ __launch_bounds__(FUSED_THREADS, fused_waves<C>()) void k_fused_pass2(
    const C *__restrict__ Ap, const uint64_t *__restrict__ itab,
    const uint64_t *__restrict__ rinfo, int nNk, int nCol, int nRow,
    const int64_t *__restrict__ ostart, const uint16_t *__restrict__ Elist,
    const C *__restrict__ dt, TO *__restrict__ y, int n, C kr, C kf,
    const double *__restrict__ scale)
{
  extern __shared__ __align__(16) unsigned char smem[];
  C *lds = reinterpret_cast<C *>(smem);
  if ((int)blockIdx.x >= nRow) return;
  const uint64_t ri = rinfo[blockIdx.x];
  const int r = (int)(ri & 0xFFFF), mr = (int)((ri >> 16) & 0xFFFF);
  const int tr = (int)((ri >> 32) & 0xFF), q = (int)((ri >> 40) & 0xFFFF);
  const uint32_t wmask = (1u << n) - 1;
  // tr mod p for every period p <= n (the mirrored items' start rotation
  // needs (m + tr) mod p with m < p)
  __shared__ int trm[32];
  if ((int)threadIdx.x <= n) trm[threadIdx.x] = threadIdx.x ? tr % (int)threadIdx.x : 0;
  __syncthreads();
  const C *Ar = Ap + (size_t)r * nCol * 16;
  const C *Am = Ap + (size_t)mr * nCol * 16;
  // item quarter t = (item, slots 4(t&3) .. 4(t&3)+3): one vector load each
  const int nt = nNk * 4;
  constexpr int UNR = 5;
  for (int t0 = threadIdx.x; t0 < nt; t0 += blockDim.x * UNR) {
    uint64_t it[UNR];
    C v[UNR][4];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int t = t0 + u * (int)blockDim.x;
      it[u] = t < nt ? itab[t >> 2] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int t = t0 + u * (int)blockDim.x;
      const int pc = (int)((it[u] >> 32) & 0xFF), k0 = 4 * (t & 3);
      if (t < nt && k0 < pc) {
        const bool mir = ((it[u] >> 40) & 0xFF) != 255;
        ld4((mir ? Am : Ar) + (size_t)(it[u] & 0xFFFF) * 16 + k0, v[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int t = t0 + u * (int)blockDim.x;
      const int pc = (int)((it[u] >> 32) & 0xFF), k0 = 4 * (t & 3);
      if (t >= nt || k0 >= pc) continue;
      const uint32_t w = (uint32_t)((it[u] >> 16) & 0xFFFF);
      const int m = (int)((it[u] >> 40) & 0xFF);
      if (m == 255) {
        // E_k = rotl_k NR[c], k = k0 .. k0+3 (k0 < pc <= n)
        uint32_t E = k0 ? (((w << k0) & wmask) | (w >> (n - k0))) : w;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (k0 + i >= pc) break;
          lds[fsw(E, n)] = v[u][i];
          E = ((E << 1) & wmask) | (E >> (n - 1));
        }
      } else {
        // E_j = rotr_{m+t+j} NR[c], j = k0 .. k0+3: d = ((m + t) mod pc + k0)
        // mod n with m < pc and k0 < pc <= n
        int d = m + trm[pc];
        if (d >= pc) d -= pc;
        d += k0;
        if (d >= n) d -= n;
        uint32_t E = d ? ((w >> d) | ((w << (n - d)) & wmask)) : w;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (k0 + i >= pc) break;
          lds[fsw(E, n)] = v[u][i];
          E = (E >> 1) | ((E & 1u) << (n - 1));
        }
      }
    }
  }
  __syncthreads();
  fused_kron(lds, n, kr);
  // 1 / scale of the input, and kf = a^n of this pass's butterflies
  const C inv = (scale ? (C)(1.0 / *scale) : (C)1) * kf;
  // internal numbering: the orbits of this row are [ostart[q], ostart[q+1]);
  // UNR index and weight loads in flight per thread before the LDS reads
  const int64_t i0 = ostart[q], i1 = ostart[q + 1];
  constexpr int UNO = 4;
  for (int64_t b = i0 + threadIdx.x; b < i1; b += (int64_t)blockDim.x * UNO) {
    uint16_t E[UNO];
    C d[UNO];
#pragma unroll
    for (int u = 0; u < UNO; ++u) {
      const int64_t i = b + (int64_t)u * blockDim.x;
      E[u] = i < i1 ? Elist[i] : (uint16_t)0;
      d[u] = i < i1 ? dt[i] : (C)0;
    }
#pragma unroll
    for (int u = 0; u < UNO; ++u) {
      const int64_t i = b + (int64_t)u * blockDim.x;
      if (i < i1) y[i] = (TO)(d[u] * lds[E[u]] * inv);
    }
  }
}

template <typename C>
void reduced_typed(const ReducedPlan &p, const void *in, void *out, tm_dtype io, void *work,
                   const double *scale, hipStream_t s, const C *tvec, const C *dt)
{
  const int64_t M = p.M;
  C *xt = static_cast<C *>(work);
  C *rest = xt + ((M + 63) / 64) * 64;
  const unsigned gM = grid_for((uint64_t)M, 256);
  const bool a16 = ((uintptr_t)in % 16) == 0 && ((uintptr_t)tvec % 16) == 0 &&
                   ((uintptr_t)xt % 16) == 0;
  const unsigned g4 = grid_for((uint64_t)(M / 4 + 1), 256);
  if (io == TM_F64 && a16)
    hipLaunchKernelGGL((k_xtilde4<C, double>), dim3(g4), dim3(256), 0, s,
                       static_cast<const double *>(in), tvec, xt, M, scale);
  else if (io == TM_F64)
    hipLaunchKernelGGL((k_xtilde<C, double>), dim3(gM), dim3(256), 0, s,
                       static_cast<const double *>(in), tvec, xt, M, scale);
  else if (a16)
    hipLaunchKernelGGL((k_xtilde4<C, float>), dim3(g4), dim3(256), 0, s,
                       static_cast<const float *>(in), tvec, xt, M, scale);
  else
    hipLaunchKernelGGL((k_xtilde<C, float>), dim3(gM), dim3(256), 0, s,
                       static_cast<const float *>(in), tvec, xt, M, scale);
  TM_LAUNCH_CHECK();
  // K = ka [[1, r], [r, 1]]: butterflies with r, the factor ka^(bits) on output
  const C kr = (C)(p.kb / p.ka);
  if (p.ft) {
    const FusedTables &ft = *p.ft;
    const int n = ft.n;
    C *Ap = rest;
    const size_t shm1 = (size_t(1) << n) * sizeof(C) + (size_t)ft.nNk * 4;
    const size_t shm2 = (size_t(1) << n) * sizeof(C);
    const C kf = dev::kron_coef<C>(p.ka, 0, p.kb, 0, n).f;
    const unsigned g1 = 8u * (unsigned)((ft.nCol + 7) / 8);
    hipLaunchKernelGGL((k_fused_pass1<C>), dim3(g1), dim3(FUSED_THREADS), shm1, s,
                       ft.jtab.as<uint32_t>(), ft.cbase.as<uint32_t>(), ft.coff.as<uint8_t>(),
                       xt, Ap, ft.slotO.as<uint16_t>(), n, ft.nCol, ft.nNk, kr, kf);
    TM_LAUNCH_CHECK();
    if (io == TM_F64)
      hipLaunchKernelGGL((k_fused_pass2<C, double>), dim3(ft.nRow), dim3(FUSED_THREADS), shm2, s,
                         Ap, ft.itab.as<uint64_t>(), ft.rinfo.as<uint64_t>(), ft.nNk, ft.nCol,
                         ft.nRow, ft.ostart.as<int64_t>(), ft.Elist.as<uint16_t>(), dt,
                         static_cast<double *>(out), n, kr, kf, scale);
    else
      hipLaunchKernelGGL((k_fused_pass2<C, float>), dim3(ft.nRow), dim3(FUSED_THREADS), shm2, s,
                         Ap, ft.itab.as<uint64_t>(), ft.rinfo.as<uint64_t>(), ft.nNk, ft.nCol,
                         ft.nRow, ft.ostart.as<int64_t>(), ft.Elist.as<uint16_t>(), dt,
                         static_cast<float *>(out), n, kr, kf, scale);
    TM_LAUNCH_CHECK();
    return;
  }
  // generic: expand, full-space operator, gather
  const uint64_t N = 1ull << p.L;
  C *u = rest;
  if (p.ylayer) {
    // symmetrised axial transfer y(n, c) = W(n, c) sum_p K2(n, p) W(c, p) x(c, p)
    // (TMat/tmatreduce.cpp:832-873); its transpose (:957-1015) in reverse order
    const int Ly = p.ylayer;
    const C *wt = sizeof(C) == 8 ? reinterpret_cast<const C *>(p.wtab64)
                                 : reinterpret_cast<const C *>(p.wtab32);
    const tm_dtype ct = sizeof(C) == 8 ? TM_F64 : TM_F32;
    hipLaunchKernelGGL((k_expand_w<C>), dim3(grid_for(N, 256)), dim3(256), 0, s, p.map, xt, u, N,
                       Ly, wt, p.ybm);
    TM_LAUNCH_CHECK();
    C *v = u + N;
    DiagSpec none{};
    none.kind = -1;
    const KronRange lo{2 * Ly, 0, Ly, false, false, none, none, nullptr, nullptr, p.ka, p.kb};
    if (!p.transpose) {
      launch_kron_range(lo, u, u, u, ct, ct, ct, nullptr, false, false, false, s);
      launch_layer_swap(u, v, ct, ct, ct, Ly, wt, 1, p.ybm, nullptr, 0, s);
    } else {
      launch_layer_swap(u, v, ct, ct, ct, Ly, wt, 2, p.ybm, nullptr, 0, s);
      launch_kron_range(lo, v, v, v, ct, ct, ct, nullptr, false, false, false, s);
    }
    // the forward operator applied W(n, c) in the swap; the transpose ends
    // with W of the output state
    if (!p.transpose && io == TM_F64)
      hipLaunchKernelGGL((k_gather<C, double>), dim3(gM), dim3(256), 0, s, v, p.reps, dt,
                         static_cast<double *>(out), M, scale);
    else if (!p.transpose)
      hipLaunchKernelGGL((k_gather<C, float>), dim3(gM), dim3(256), 0, s, v, p.reps, dt,
                         static_cast<float *>(out), M, scale);
    else if (io == TM_F64)
      hipLaunchKernelGGL((k_gather_w<C, double>), dim3(gM), dim3(256), 0, s, v, p.reps, dt,
                         static_cast<double *>(out), M, scale, Ly, wt, p.ybm);
    else
      hipLaunchKernelGGL((k_gather_w<C, float>), dim3(gM), dim3(256), 0, s, v, p.reps, dt,
                         static_cast<float *>(out), M, scale, Ly, wt, p.ybm);
    TM_LAUNCH_CHECK();
    return;
  }
  hipLaunchKernelGGL((k_expand<C>), dim3(grid_for(N, 256)), dim3(256), 0, s, p.map, xt, u, N);
  TM_LAUNCH_CHECK();
  const int smallL = sizeof(C) == 8 ? 13 : 14;
  if (p.sweep) {
    // gate-circuit operator (XDT or ZBT) with its own diagonal and normfactor
    C *v = u + N;
    const tm_dtype ct = sizeof(C) == 8 ? TM_F64 : TM_F32;
    launch_sweep(*p.sweep, u, v, ct, ct, p.transpose, v + N, nullptr, s);
    u = v;
  } else if (p.L <= smallL) {
    hipLaunchKernelGGL((k_kron_small<C>), dim3(1), dim3(1024), (size_t(1) << p.L) * sizeof(C),
                       s, u, p.L, kr, dev::kron_coef<C>(p.ka, 0, p.kb, 0, p.L).f);
    TM_LAUNCH_CHECK();
  } else {
    KronPlan kp{};
    kp.L = p.L;
    kp.prologue = kp.epilogue = false;
    kp.ka = p.ka;
    kp.kb = p.kb;
    kp.diag.kind = -1;
    const tm_dtype ct = sizeof(C) == 8 ? TM_F64 : TM_F32;
    launch_kron(kp, u, u, ct, ct, ct, nullptr, nullptr, s);
  }
  if (io == TM_F64)
    hipLaunchKernelGGL((k_gather<C, double>), dim3(gM), dim3(256), 0, s, u, p.reps, dt,
                       static_cast<double *>(out), M, scale);
  else
    hipLaunchKernelGGL((k_gather<C, float>), dim3(gM), dim3(256), 0, s, u, p.reps, dt,
                       static_cast<float *>(out), M, scale);
  TM_LAUNCH_CHECK();
}

}  // namespace

size_t reduced_work_bytes(const ReducedPlan &p, tm_dtype comp)
{
  const size_t es = comp == TM_F64 ? 8 : 4;
  size_t b = (size_t)((p.M + 63) / 64) * 64 * es;
  if (p.ft) {
    b += (size_t)p.ft->nCol * p.ft->nNk * 16 * es;
  } else if (p.sweep) {
    b += 2 * (size_t(1) << p.L) * es + sweep_work_bytes(p.sweep->kind, p.L, comp);
  } else if (p.ylayer) {
    b += 2 * (size_t(1) << p.L) * es;
  } else {
    b += (size_t(1) << p.L) * es;
  }
  return b;
}

void launch_reduced(const ReducedPlan &p, const void *in, void *out, tm_dtype io, tm_dtype comp,
                    void *work, const double *scale, hipStream_t s)
{
  if (comp == TM_F64)
    reduced_typed<double>(p, in, out, io, work, scale, s, p.tvec, p.dt64);
  else
    reduced_typed<float>(p, in, out, io, work, scale, s, p.tvec32, p.dt32);
}

namespace {
template <typename T>
__global__ void k_perm_gather(const T *__restrict__ x, const uint32_t *__restrict__ perm,
                              T *__restrict__ xi, int64_t M)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M;
       i += (int64_t)gridDim.x * blockDim.x)
    xi[i] = x[perm[i]];
}
template <typename T>
__global__ void k_perm_scatter(const T *__restrict__ yi, const uint32_t *__restrict__ perm,
                               T *__restrict__ y, int64_t M)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < M;
       i += (int64_t)gridDim.x * blockDim.x)
    y[perm[i]] = yi[i];
}
}  // namespace

void permute_to_internal(const void *x, void *xi, const uint32_t *perm, int64_t M, tm_dtype io,
                         hipStream_t s)
{
  const unsigned g = grid_for((uint64_t)M, 256);
  if (io == TM_F64)
    hipLaunchKernelGGL(k_perm_gather<double>, dim3(g), dim3(256), 0, s,
                       static_cast<const double *>(x), perm, static_cast<double *>(xi), M);
  else
    hipLaunchKernelGGL(k_perm_gather<float>, dim3(g), dim3(256), 0, s,
                       static_cast<const float *>(x), perm, static_cast<float *>(xi), M);
  TM_LAUNCH_CHECK();
}

void permute_to_external(const void *yi, void *y, const uint32_t *perm, int64_t M, tm_dtype io,
                         hipStream_t s)
{
  const unsigned g = grid_for((uint64_t)M, 256);
  if (io == TM_F64)
    hipLaunchKernelGGL(k_perm_scatter<double>, dim3(g), dim3(256), 0, s,
                       static_cast<const double *>(yi), perm, static_cast<double *>(y), M);
  else
    hipLaunchKernelGGL(k_perm_scatter<float>, dim3(g), dim3(256), 0, s,
                       static_cast<const float *>(yi), perm, static_cast<float *>(y), M);
  TM_LAUNCH_CHECK();
}

}  // namespace ltm
