// Learned candidate ranking for point location over a partition (no counterpart
// in the reference; an addition on top of the quadtree join).
//
// A coarse grid (dim x dim cells over the polygons' domain) lists, per cell, up to
// 16 candidate polygons: every polygon whose bounding box reaches the cell, so the
// list of a point's cell holds every polygon that can contain the point. A small
// MLP scores the candidates of a point's cell from the cell's learned embedding
// (64 values) and 32 Fourier features of the point's position inside the cell:
//
//     x[96] = [E[cell][0..63], sin/cos(pi k u), sin/cos(pi k v), k = 1..8]
//     h1 = relu(W1 x + b1)  (128)    h2 = relu(W2 h1 + b2)  (128)
//     logits = W3 h2 + b3   (16, one per candidate slot)
//
// bf16 operands, fp32 accumulation, on MFMA (v_mfma_f32_16x16x32_bf16 for the
// locator, v_mfma_f32_16x16x16_bf16 in the training GEMM). The
// locator tests the candidates in the order of their logits with the exact FP64
// point-in-polygon test of the quadtree path (same edge table, same arithmetic)
// and stops at the first polygon that contains the point away from its edges.
// Points whose only hit is close to an edge, or whose cell lists more than 16
// candidates, or that have no hit after max_tests tests, are marked for the
// exact quadtree path; everything else is final.
//
// Training runs forward and backward passes on the same MFMA tiles (a generic
// 16x16-tile GEMM with strided operands) plus small kernels for the masked
// softmax cross-entropy, bias and embedding gradients and Adam.
#include <algorithm>
#include <cmath>

#include "qpip_edges.hpp"

namespace qpip {

namespace {

constexpr int kCand   = 16;  // candidate slots per grid cell
constexpr int kEmb    = 64;  // cell embedding width
constexpr int kFour   = 32;  // Fourier features
constexpr int kIn     = 96;  // kEmb + kFour
constexpr int kHid    = 128; // hidden width
constexpr int kFreqs  = 8;   // frequencies per axis (kFour = 2 axes x 8 x sin/cos)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round to nearest even (v_cvt_pk_bf16_f32); NaN stays NaN.
__device__ __forceinline__ uint16_t to_bf16(float f)
{
  return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}
__device__ __forceinline__ float from_bf16(uint16_t h)
{
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

// D = A B + C on one 16x16x16 bf16 tile. Lane l supplies A[l % 16][4 (l / 16) + j]
// and B[4 (l / 16) + j][l % 16] (j = 0..3) and holds D[4 (l / 16) + j][l % 16].
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Flat parameter layout (floats; the bf16 copy uses the same element offsets).
struct Layout {
  int64_t e, w1, b1, w2, b2, w3, b3, total;
  __host__ __device__ Layout() : Layout(0) {}
  __host__ __device__ explicit Layout(int64_t cells)
  {
    e     = 0;
    w1    = cells * kEmb;
    b1    = w1 + int64_t{kHid} * kIn;
    w2    = b1 + kHid;
    b2    = w2 + int64_t{kHid} * kHid;
    w3    = b2 + kHid;
    b3    = w3 + int64_t{kCand} * kHid;
    total = b3 + kCand;
  }
};

struct GridView {
  double x0, y0, cell, inv_cell;
  uint32_t dim;
  int32_t const* cand;
  uint8_t const* count;
};

// Cell coordinate floor((v - v0) * inv_cell) with inv_cell = 1 / cell, clamped to
// [0, dim - 1] (NaN -> 0). The function is monotone in v, so a polygon whose box
// spans [lo, hi] is listed in every cell from coord(lo) to coord(hi), which
// includes the cell of each point it contains.
__device__ __forceinline__ uint32_t grid_coord(double v, double v0, double inv_cell, uint32_t dim)
{
  double const t = floor((v - v0) * inv_cell);
  if (!(t >= 0.0)) return 0u;
  if (t >= static_cast<double>(dim - 1)) return dim - 1;
  return static_cast<uint32_t>(t);
}

struct Loc {
  uint32_t cell;
  float u, v;
};

__device__ __forceinline__ Loc locate_cell(GridView const& g, double x, double y)
{
  uint32_t const ix = grid_coord(x, g.x0, g.inv_cell, g.dim);
  uint32_t const iy = grid_coord(y, g.y0, g.inv_cell, g.dim);
  double const u    = (x - (g.x0 + static_cast<double>(ix) * g.cell)) * g.inv_cell;
  double const v    = (y - (g.y0 + static_cast<double>(iy) * g.cell)) * g.inv_cell;
  return Loc{iy * g.dim + ix, static_cast<float>(u), static_cast<float>(v)};
}

// Fourier features of frequency index k (0..7): sin, cos of pi (k+1) u, then of v,
// from the hardware sine and cosine (v_sin_f32 / v_cos_f32 take revolutions:
// sin(pi (k+1) u) = sin(2 pi r) with r = (k+1) u / 2). Training and location use
// this same function, so the model sees identical features in both.
__device__ __forceinline__ void fourier4(float u, float v, int k, float (&f)[4])
{
  float const h  = 0.5f * static_cast<float>(k + 1);
  float const ru = h * u, rv = h * v;
  f[0]           = __builtin_amdgcn_sinf(ru);
  f[1]           = __builtin_amdgcn_cosf(ru);
  f[2]           = __builtin_amdgcn_sinf(rv);
  f[3]           = __builtin_amdgcn_cosf(rv);
}

// ---------------------------------------------------------------------------------
// Training kernels
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
  features_kernel(GridView g, uint16_t const* __restrict__ emb, double const* __restrict__ xy,
                  int64_t n, int64_t rows, uint16_t* __restrict__ X, uint32_t* __restrict__ cells)
{
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  uint16_t* x = X + i * kIn;
  if (i >= n) {  // padding rows: zero features
    for (int d = 0; d < kIn; ++d) x[d] = 0;
    cells[i] = 0;
    return;
  }
  Loc const L = locate_cell(g, xy[2 * i], xy[2 * i + 1]);
  cells[i]    = L.cell;
  uint16_t const* e = emb + static_cast<int64_t>(L.cell) * kEmb;
  for (int d = 0; d < kEmb; ++d) x[d] = e[d];
  for (int k = 0; k < kFreqs; ++k) {
    float f[4];
    fourier4(L.u, L.v, k, f);
    for (int j = 0; j < 4; ++j) x[kEmb + 4 * k + j] = to_bf16(f[j]);
  }
}

// C(m, n) = sum_k A(m, k) B(k, n) over 16x16 tiles, one wave per (tile, K chunk).
// A(m, k) = A[m a_m + k a_k], B(k, n) = B[k b_k + n b_n] (bf16). M and N are
// multiples of 16. atomic: C += partial (fp32) for split-K; otherwise the
// epilogue adds bias[n], applies relu and the gate (keep where gate(m, n) > 0)
// and stores fp32 C and/or bf16 Cb.
struct GemmArgs {
  uint16_t const* A;
  int64_t a_m, a_k;
  uint16_t const* B;
  int64_t b_k, b_n;
  int64_t M, N, K, kchunk;
  float* C;
  int64_t c_m, c_n;
  uint16_t* Cb;
  int64_t cb_m, cb_n;
  float const* bias;
  uint16_t const* gate;
  int64_t g_m, g_n;
  int relu, atomic;
};

__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs a)
{
  int64_t const w      = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
  int64_t const tiles_n = a.N / 16;
  int64_t const chunks  = (a.K + a.kchunk - 1) / a.kchunk;
  if (w >= (a.M / 16) * tiles_n * chunks) return;
  int64_t const chunk = w % chunks;
  int64_t const t     = w / chunks;
  int64_t const tm = t / tiles_n, tn = t % tiles_n;
  int const l = lane_id(), g = l >> 4, r = l & 15;
  int64_t const am = tm * 16 + r;
  int64_t const bn = tn * 16 + r;
  int64_t const k0 = chunk * a.kchunk;
  int64_t const k1 = k0 + a.kchunk < a.K ? k0 + a.kchunk : a.K;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int64_t k = k0; k < k1; k += 16) {
    bf16x4 av, bv;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int64_t const kk = k + 4 * g + j;
      bool const in    = kk < k1;
      av[j] = in ? static_cast<short>(a.A[am * a.a_m + kk * a.a_k]) : short{0};
      bv[j] = in ? static_cast<short>(a.B[kk * a.b_k + bn * a.b_n]) : short{0};
    }
    acc = mfma16(av, bv, acc);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int64_t const m = tm * 16 + 4 * g + j;
    int64_t const n = tn * 16 + r;
    float v         = acc[j];
    if (a.atomic) {
      atomicAdd(a.C + m * a.c_m + n * a.c_n, v);
      continue;
    }
    if (a.bias) v += a.bias[n];
    if (a.relu) v = fmaxf(v, 0.f);
    if (a.gate && !(from_bf16(a.gate[m * a.g_m + n * a.g_n]) > 0.f)) v = 0.f;
    if (a.C) a.C[m * a.c_m + n * a.c_n] = v;
    if (a.Cb) a.Cb[m * a.cb_m + n * a.cb_n] = to_bf16(v);
  }
}

// Masked softmax cross-entropy over the first count[cell] slots. Rows with
// label < 0 (or outside the mask) contribute nothing. dZ = (p - onehot) * scale.
// stats += {loss sum, correct top-1, labelled rows}.
__global__ void __launch_bounds__(256)
  xent_kernel(float const* __restrict__ Z, uint32_t const* __restrict__ cells,
              uint8_t const* __restrict__ count, int32_t const* __restrict__ labels, int64_t n,
              int64_t rows, float scale, uint16_t* __restrict__ dZ, float* __restrict__ stats)
{
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  int const c   = i < n ? (count[cells[i]] < kCand ? count[cells[i]] : kCand) : 0;
  int const y   = i < n ? labels[i] : -1;
  float const* z = Z + i * kCand;
  if (y < 0 || y >= c) {
    for (int s = 0; s < kCand; ++s) dZ[i * kCand + s] = 0;
    return;
  }
  float mx = -INFINITY;
  int best = 0;
  for (int s = 0; s < c; ++s)
    if (z[s] > mx) {
      mx   = z[s];
      best = s;
    }
  float sum = 0.f;
  for (int s = 0; s < c; ++s) sum += expf(z[s] - mx);
  for (int s = 0; s < kCand; ++s) {
    float const p = s < c ? expf(z[s] - mx) / sum : 0.f;
    dZ[i * kCand + s] = to_bf16((p - (s == y ? 1.f : 0.f)) * scale);
  }
  atomicAdd(stats + 0, logf(sum) - (z[y] - mx));
  if (best == y) atomicAdd(stats + 1, 1.f);
  atomicAdd(stats + 2, 1.f);
}

// out[c] += sum over rows of D[row][c] (bf16, `cols` columns, a divisor of 256).
// Thread t reads column t % cols of rows t / cols, t / cols + 256 / cols, ...
// (consecutive threads read consecutive bf16 of a row: coalesced), 16 rows each;
// the block's partial sums per column meet in LDS, one atomic per column.
constexpr int kColsumIters = 16;
inline int64_t colsum_rows_per_block(int cols) { return int64_t{256} / cols * kColsumIters; }

__global__ void __launch_bounds__(256)
  colsum_kernel(uint16_t const* __restrict__ D, int64_t rows, int cols, float* __restrict__ out)
{
  __shared__ float part[256];
  int const c            = threadIdx.x % cols;
  int64_t const step     = 256 / cols;
  int64_t const r0       = static_cast<int64_t>(blockIdx.x) * step * kColsumIters;
  int64_t const r_end    = r0 + step * kColsumIters < rows ? r0 + step * kColsumIters : rows;
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kColsumIters; ++k) {
    int64_t const r = r0 + threadIdx.x / cols + k * step;
    if (r < r_end) s += from_bf16(D[r * cols + c]);
  }
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < cols) {
    float t = 0.f;
    for (int k = threadIdx.x; k < 256; k += cols) t += part[k];
    atomicAdd(out + threadIdx.x, t);
  }
}

// gE[cell_i][d] += dXe[i][d]
__global__ void __launch_bounds__(256)
  embed_grad_kernel(float const* __restrict__ dXe, uint32_t const* __restrict__ cells, int64_t n,
                    float* __restrict__ gE)
{
  int64_t const t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n * kEmb) return;
  int64_t const i = t / kEmb;
  int const d     = static_cast<int>(t % kEmb);
  atomicAdd(gE + static_cast<int64_t>(cells[i]) * kEmb + d, dXe[t]);
}

__global__ void __launch_bounds__(256)
  adam_kernel(float* __restrict__ p, uint16_t* __restrict__ pb, float const* __restrict__ grad,
              float* __restrict__ m, float* __restrict__ v, int64_t n, float lr, float b1, float b2,
              float eps, float c1, float c2)
{
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float const g  = grad[i];
  float const mi = b1 * m[i] + (1.f - b1) * g;
  float const vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i]           = mi;
  v[i]           = vi;
  float const q  = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  p[i]           = q;
  pb[i]          = to_bf16(q);
}

__global__ void __launch_bounds__(256) fill_f32_kernel(float* __restrict__ p, int64_t n, float v)
{
  int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------------
// Locator: MLP on MFMA + exact test in predicted order
// ---------------------------------------------------------------------------------
constexpr int kLocTiles   = 4;  // 16-point tiles per wave round: 64 points, one per lane
constexpr int kKT         = 32;  // K per v_mfma_f32_16x16x32_bf16
constexpr int kW1Frags    = (kHid / 16) * (kIn / kKT);    // 24
constexpr int kW2Frags    = (kHid / 16) * (kHid / kKT);   // 32
constexpr int kW3Frags    = (kCand / 16) * (kHid / kKT);  // 4
constexpr int kFrags      = kW1Frags + kW2Frags + kW3Frags;
constexpr int kFragElems  = 8;  // bf16 per lane per fragment
constexpr int kLogitPitch = kCand + 1;  // LDS row pitch of the logit transpose
constexpr int kLocWaves   = 4;

struct PolyView {
  uint32_t const* poly_offsets;
  int64_t num_polys;
  uint32_t const* ring_offsets;
  int64_t num_rings;
  Edge const* edges;
  EdgeB const* edges_b;
  uint32_t const* ring_edges;
  struct Record const* records;
};

// Per polygon with exactly one ring: its effective edges [e0, e0 + ne) and the y
// the ring's walk starts from (edges_b[e0].by), so a test needs one load before
// the edges. ne == kMultiRing marks a polygon tested through the ring tables.
constexpr uint32_t kMultiRing = 0xffffffffu;
struct alignas(16) Record {
  uint32_t e0, ne;
  double y0;
};

__global__ void __launch_bounds__(256) records_kernel(PolyView P, Record* __restrict__ out)
{
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < P.num_polys;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    [&] {
      uint32_t const r0 = P.poly_offsets[i];
      uint32_t r1       = P.poly_offsets[i + 1];
      if (r1 > P.num_rings) r1 = static_cast<uint32_t>(P.num_rings);
      Record rec{0u, kMultiRing, 0.0};
      if (r1 == r0 + 1) {
        rec.e0 = P.ring_offsets[r0];
        rec.ne = P.ring_edges[r0];
        rec.y0 = rec.ne ? P.edges_b[rec.e0].by : 0.0;
      } else if (r1 <= r0) {
        rec.ne = 0;  // no rings: never contains a point
      }
      out[i] = rec;
    }();
}

struct LocateArgs {
  GridView g;
  uint16_t const* pb;  // bf16 parameters
  float const* p;      // fp32 parameters (biases)
  Layout lay;
  PolyView P;
  double const* xy;
  int64_t n;
  int max_tests;
  int32_t* out_poly;
  uint8_t* out_tests;
  uint8_t* out_flags;
  uint16_t* order;  // per point: the best 3 slots (4 bits each), then their number
};

// The exact test of the quadtree path for one point and one polygon, plus whether
// the point lies within the collinearity filter of any edge (then a neighbouring
// polygon's test could also hold, and the point goes to the quadtree path).
__device__ __forceinline__ void edge_step(Edge const& E, EdgeB const* eb, double px, double py,
                                          bool& y0, bool& inside, bool& on_edge, bool& near)
{
  double const lhs = (px - E.ax) * E.rise;  // run_to_point * rise
  double const rhs = E.run * (py - E.ay);   // run * rise_to_point
  double const gap = fma(fabs(rhs), -0x1p-48, fabs(rhs - lhs));
  if (!(gap > 0x1p-1020)) {
    near = true;
    if (float_equal(rhs, lhs)) {  // is_point_in_polygon.cuh:56-65
      double const bx   = eb->bx;
      double const minx = E.ax < bx ? E.ax : bx;
      double const maxx = E.ax < bx ? bx : E.ax;
      if (minx <= px && px <= maxx) on_edge = true;
    }
  }
  bool const y1 = E.ay > py;
  if (y1 != y0 && ((lhs < rhs) != y1)) inside = !inside;
  y0 = y1;
}

__device__ __forceinline__ void exact_test(PolyView const& P, uint32_t poly, double px, double py,
                                           bool& inside_out, bool& near_out)
{
  Record const rec = P.records[poly];
  if (rec.ne != kMultiRing) {
    bool inside = false, on_edge = false, near = false;
    bool y0 = rec.y0 > py;
    // four edges' loads in flight per wait (a lane's edges are its own, so each
    // load is a round trip to L2; vector loads complete in order)
    Edge const* E = P.edges + rec.e0;
    EdgeB const* B = P.edges_b + rec.e0;
    uint32_t e = 0;
    for (; e + 4 <= rec.ne; e += 4) {
      Edge const e0_ = E[e], e1_ = E[e + 1], e2_ = E[e + 2], e3_ = E[e + 3];
      edge_step(e0_, B + e, px, py, y0, inside, on_edge, near);
      edge_step(e1_, B + e + 1, px, py, y0, inside, on_edge, near);
      edge_step(e2_, B + e + 2, px, py, y0, inside, on_edge, near);
      edge_step(e3_, B + e + 3, px, py, y0, inside, on_edge, near);
    }
    for (; e < rec.ne; ++e) edge_step(E[e], B + e, px, py, y0, inside, on_edge, near);
    inside_out = inside && !on_edge;
    near_out   = near;
    return;
  }
  uint32_t const r0 = P.poly_offsets[poly];
  uint32_t r1       = P.poly_offsets[poly + 1];
  if (r1 > P.num_rings) r1 = static_cast<uint32_t>(P.num_rings);
  bool inside = false, on_edge = false, near = false;
  for (uint32_t r = r0; r < r1; ++r) {
    uint32_t const e0 = P.ring_offsets[r];
    uint32_t const ne = P.ring_edges[r];
    if (ne == 0) continue;
    bool y0 = P.edges_b[e0].by > py;
    for (uint32_t e = 0; e < ne; ++e)
      edge_step(P.edges[e0 + e], P.edges_b + e0 + e, px, py, y0, inside, on_edge, near);
  }
  inside_out = inside && !on_edge;
  near_out   = near;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// D = A B + C on one 16x16x32 bf16 tile (gfx950). Lane l supplies
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16] (j = 0..7) and holds
// D[4 (l / 16) + j][l % 16] (j = 0..3).
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 load_frag(uint16_t const* lds, int frag, int lane)
{
  return *reinterpret_cast<bf16x8 const*>(lds +
                                          (static_cast<int64_t>(frag) * kWave + lane) * kFragElems);
}

// Input column of k-slot j of K tile c for lane group g of a layer fed by the
// previous layer's output tiles: the B operand of that tile is output tiles 2c
// (slots 0..3: rows 4g + j) and 2c + 1 (slots 4..7: rows 16 + 4g + j - 4), so the
// weights of slot j sit in that column. The first layer's input is in natural
// order (32c + 8g + j).
__host__ __device__ __forceinline__ int chained_col(int c, int g, int j)
{
  return 32 * c + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

__device__ __forceinline__ __bf16 bf(float f) { return static_cast<__bf16>(f); }

// Two waves per SIMD (at most 256 registers, which also keeps the accumulators in
// VGPRs: no AGPR copies before the bf16 conversions), as many as two workgroups'
// LDS allows: one wave's feature and conversion work overlaps the other's MFMAs.
__global__ void __launch_bounds__(kLocWaves * kWave, 2) rank_kernel(LocateArgs a)
{
  __shared__ __attribute__((aligned(16))) uint16_t w_lds[kFrags * kWave * kFragElems];
  __shared__ float b_lds[kHid + kHid + kCand];
  __shared__ float logit_lds[kLocWaves][kWave * kLogitPitch];

  // Stage W1, W2, W3 in fragment order: fragment f (M tile mt, K tile c), lane l,
  // element j holds W[16 mt + l % 16][column of slot j of lane group l / 16].
  for (int t = threadIdx.x; t < kFrags * kWave; t += blockDim.x) {
    int const f = t / kWave, l = t % kWave, g = l >> 4;
    uint16_t const* W;
    int in, mt, c;
    bool chained;
    if (f < kW1Frags) {
      W = a.pb + a.lay.w1, in = kIn, mt = f / (kIn / kKT), c = f % (kIn / kKT), chained = false;
    } else if (f < kW1Frags + kW2Frags) {
      int const ff = f - kW1Frags;
      W = a.pb + a.lay.w2, in = kHid, mt = ff / (kHid / kKT), c = ff % (kHid / kKT), chained = true;
    } else {
      int const ff = f - kW1Frags - kW2Frags;
      W = a.pb + a.lay.w3, in = kHid, mt = ff / (kHid / kKT), c = ff % (kHid / kKT), chained = true;
    }
    uint16_t const* row = W + static_cast<int64_t>(16 * mt + (l & 15)) * in;
#pragma unroll
    for (int j = 0; j < kFragElems; ++j)
      w_lds[t * kFragElems + j] = row[chained ? chained_col(c, g, j) : kKT * c + 8 * g + j];
  }
  for (int t = threadIdx.x; t < kHid + kHid + kCand; t += blockDim.x)
    b_lds[t] = t < kHid ? a.p[a.lay.b1 + t]
                        : (t < 2 * kHid ? a.p[a.lay.b2 + t - kHid] : a.p[a.lay.b3 + t - 2 * kHid]);
  __syncthreads();

  int const wave = threadIdx.x / kWave;
  int const l = lane_id(), g = l >> 4, r = l & 15;
  float* lg   = logit_lds[wave];
  int64_t const waves = static_cast<int64_t>(gridDim.x) * kLocWaves;
  uint16_t const* W1s = w_lds;
  uint16_t const* W2s = w_lds + kW1Frags * kWave * kFragElems;
  uint16_t const* W3s = w_lds + (kW1Frags + kW2Frags) * kWave * kFragElems;

  for (int64_t base = (static_cast<int64_t>(blockIdx.x) * kLocWaves + wave) * kWave; base < a.n;
       base += waves * kWave) {
    uint32_t cell[kLocTiles];
    bf16x8 X[kLocTiles][kIn / kKT];
#pragma unroll
    for (int t = 0; t < kLocTiles; ++t) {
      int64_t const p = base + 16 * t + r;
      bool const in   = p < a.n;
      Loc const L     = locate_cell(a.g, in ? a.xy[2 * p] : 0.0, in ? a.xy[2 * p + 1] : 0.0);
      cell[t]         = L.cell;
      uint16_t const* e = a.pb + a.lay.e + static_cast<int64_t>(L.cell) * kEmb;
#pragma unroll
      for (int c = 0; c < kEmb / kKT; ++c)
        X[t][c] = *reinterpret_cast<bf16x8 const*>(e + kKT * c + 8 * g);
      // features 64 + 8g + j: frequencies 2g (j < 4) and 2g + 1 (j >= 4)
      float f0[4], f1[4];
      fourier4(L.u, L.v, 2 * g, f0);
      fourier4(L.u, L.v, 2 * g + 1, f1);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        X[t][kEmb / kKT][j]     = bf(f0[j]);
        X[t][kEmb / kKT][j + 4] = bf(f1[j]);
      }
    }
    // layer 1: output tile mt holds rows 16 mt + 4g + j; tiles 2c and 2c + 1 form
    // layer 2's K tile c (chained_col)
    f32x4 H[kLocTiles][kHid / 16];
#pragma unroll
    for (int mt = 0; mt < kHid / 16; ++mt) {
#pragma unroll
      for (int t = 0; t < kLocTiles; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) H[t][mt][j] = b_lds[16 * mt + 4 * g + j];
#pragma unroll
      for (int c = 0; c < kIn / kKT; ++c) {
        bf16x8 const w = load_frag(W1s, mt * (kIn / kKT) + c, l);
#pragma unroll
        for (int t = 0; t < kLocTiles; ++t) H[t][mt] = mfma32(w, X[t][c], H[t][mt]);
      }
    }
    // ReLU after the conversion: a bf16 value is negative exactly when its 16 bits
    // are as a signed integer, so a packed int16 max with 0 clamps two values per
    // instruction (bf16(max(x, 0)) and max(bf16(x), +0) agree for every non-NaN x)
    auto relu2 = [](float x, float y) {  // two values -> one word of bf16 pairs
      s16x2 v = __builtin_bit_cast(s16x2, __builtin_convertvector(f32x2{x, y}, bf16x2));
      return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, s16x2{0, 0}));
    };
    auto chain = [&](f32x4 const (&h)[kHid / 16], int c) {
      u32x4 const w{relu2(h[2 * c][0], h[2 * c][1]), relu2(h[2 * c][2], h[2 * c][3]),
                    relu2(h[2 * c + 1][0], h[2 * c + 1][1]),
                    relu2(h[2 * c + 1][2], h[2 * c + 1][3])};
      return __builtin_bit_cast(bf16x8, w);
    };
    // layer 2
    f32x4 H2[kLocTiles][kHid / 16];
    {
      bf16x8 B1[kLocTiles][kHid / kKT];
#pragma unroll
      for (int t = 0; t < kLocTiles; ++t)
#pragma unroll
        for (int c = 0; c < kHid / kKT; ++c) B1[t][c] = chain(H[t], c);
#pragma unroll
      for (int mt = 0; mt < kHid / 16; ++mt) {
#pragma unroll
        for (int t = 0; t < kLocTiles; ++t)
#pragma unroll
          for (int j = 0; j < 4; ++j) H2[t][mt][j] = b_lds[kHid + 16 * mt + 4 * g + j];
#pragma unroll
        for (int c = 0; c < kHid / kKT; ++c) {
          bf16x8 const w = load_frag(W2s, mt * (kHid / kKT) + c, l);
#pragma unroll
          for (int t = 0; t < kLocTiles; ++t) H2[t][mt] = mfma32(w, B1[t][c], H2[t][mt]);
        }
      }
    }
    // layer 3: logits, then transpose through LDS (lane l -> point base + l)
    {
      f32x4 acc[kLocTiles];
#pragma unroll
      for (int t = 0; t < kLocTiles; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[t][j] = b_lds[2 * kHid + 4 * g + j];
#pragma unroll
      for (int c = 0; c < kHid / kKT; ++c) {
        bf16x8 const w = load_frag(W3s, c, l);
#pragma unroll
        for (int t = 0; t < kLocTiles; ++t) acc[t] = mfma32(w, chain(H2[t], c), acc[t]);
      }
#pragma unroll
      for (int t = 0; t < kLocTiles; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) lg[(16 * t + r) * kLogitPitch + 4 * g + j] = acc[t][j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float z[kCand];
#pragma unroll
    for (int s = 0; s < kCand; ++s) z[s] = lg[l * kLogitPitch + s];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    // this lane's point is tile g, row r: base + 16 g + r = base + l
    int64_t const p = base + l;
    if (p >= a.n) continue;
    uint32_t c = cell[0];
#pragma unroll
    for (int t = 1; t < kLocTiles; ++t)
      if (g == t) c = cell[t];
    int const cnt   = a.g.count[c];
    int const limit = cnt > kCand ? 0 : (cnt < a.max_tests ? cnt : a.max_tests);
    // The three best slots by decreasing logit (NaN lowest, near-ties to the lower
    // slot): each slot's key is the logit's order-preserving bits with the low 4
    // bits replaced by 15 - slot, so keys are distinct and three max passes find
    // them. The other slots are tested in list order after these; the order only
    // decides how many exact tests a point needs, never the answer.
    uint32_t key[kCand];
#pragma unroll
    for (int q = 0; q < kCand; ++q) {
      float const zq   = z[q] == z[q] ? z[q] : -INFINITY;
      uint32_t const b = __float_as_uint(zq);
      uint32_t const o = (b >> 31) ? ~b : (b | 0x80000000u);
      key[q]           = q < cnt ? ((o & 0xFFFFFFF0u) | static_cast<uint32_t>(15 - q)) : 0u;
    }
    uint32_t m0 = 0, m1 = 0, m2 = 0;
#pragma unroll
    for (int q = 0; q < kCand; ++q) m0 = max(m0, key[q]);
#pragma unroll
    for (int q = 0; q < kCand; ++q) m1 = max(m1, key[q] == m0 ? 0u : key[q]);
#pragma unroll
    for (int q = 0; q < kCand; ++q) m2 = max(m2, (key[q] == m0 || key[q] == m1) ? 0u : key[q]);
    uint32_t const ranked = static_cast<uint32_t>(limit < 3 ? limit : 3);
    a.order[p] = static_cast<uint16_t>((15u - (m0 & 15u)) | ((15u - (m1 & 15u)) << 4) |
                                       ((15u - (m2 & 15u)) << 8) | (ranked << 12));
  }
}

// Exact tests in predicted order, one lane per point.
__global__ void __launch_bounds__(256) verify_kernel(LocateArgs a)
{
  int64_t const p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  double const x  = a.xy[2 * p];
  double const y  = a.xy[2 * p + 1];
  uint32_t const order = a.order[p];  // issued with the point, before the cell lookup
  uint32_t const c = locate_cell(a.g, x, y).cell;
  int const cnt    = a.g.count[c];
  int32_t result   = -1;
  uint8_t flags    = 0;
  int tests        = 0;
  if (cnt > kCand) {
    flags = 1;  // overflowed cell: quadtree path
  } else {
    uint32_t const ranked = order >> 12;
    int const limit       = cnt < a.max_tests ? cnt : a.max_tests;
    // one candidate: true when the search is over (a hit, or an error)
    auto test = [&](int s) {
      ++tests;
      int32_t const poly = a.g.cand[static_cast<int64_t>(c) * kCand + s];
      if (poly < 0 || poly >= a.P.num_polys) {
        flags |= 1;
        return true;
      }
      bool inside, near;
      exact_test(a.P, static_cast<uint32_t>(poly), x, y, inside, near);
      if (!inside) return false;
      if (near)
        flags |= 1;
      else
        result = poly;
      return true;
    };
    bool done = false;
    // the ranked slots, then the others in list order
    for (uint32_t i = 0; i < ranked && !done && tests < limit; ++i)
      done = test(static_cast<int>((order >> (4 * i)) & 15u));
    for (int q = 0; q < cnt && !done && tests < limit; ++q) {
      bool ranked_q = false;
      for (uint32_t i = 0; i < ranked; ++i)
        ranked_q |= ((order >> (4 * i)) & 15u) == static_cast<uint32_t>(q);
      if (!ranked_q) done = test(q);
    }
    if (result < 0 && !(flags & 1) && tests < cnt) flags |= 1;  // stopped at max_tests
    if (!(result >= 0 && tests == 1)) flags |= 2;  // first prediction was not the answer
  }
  a.out_poly[p]  = (flags & 1) ? -2 : result;
  a.out_tests[p] = static_cast<uint8_t>(tests);
  a.out_flags[p] = flags;
}

// ---------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------
GridView grid_view(qpip_grid const* g)
{
  QPIP_EXPECTS(g != nullptr, "grid must not be null");
  QPIP_EXPECTS(g->dim > 0 && g->dim <= 65536, "grid dimension must be in [1, 65536]");
  QPIP_EXPECTS(g->cell > 0.0, "grid cell size must be positive");
  return GridView{g->x0, g->y0, g->cell, 1.0 / g->cell, static_cast<uint32_t>(g->dim),
                  g->candidates, g->counts};
}

void check_model(qpip_ranker const* m, qpip_grid const* g)
{
  QPIP_EXPECTS(m != nullptr && m->params != nullptr && m->params_bf16 != nullptr,
               "ranker parameters must not be null");
  QPIP_EXPECTS(m->num_cells == int64_t{g->dim} * g->dim,
               "ranker embedding count must equal the grid's cell count");
}

void launch_gemm(GemmArgs const& a, hipStream_t stream)
{
  int64_t const waves = (a.M / 16) * (a.N / 16) * ((a.K + a.kchunk - 1) / a.kchunk);
  if (waves == 0) return;
  gemm_kernel<<<grid_for(static_cast<uint64_t>(waves), 256 / kWave, 1u << 30), 256, 0, stream>>>(a);
  QPIP_LAUNCH_CHECK();
}

GemmArgs gemm(uint16_t const* A, int64_t a_m, int64_t a_k, uint16_t const* B, int64_t b_k,
              int64_t b_n, int64_t M, int64_t N, int64_t K)
{
  GemmArgs a{};
  a.A = A, a.a_m = a_m, a.a_k = a_k, a.B = B, a.b_k = b_k, a.b_n = b_n;
  a.M = M, a.N = N, a.K = K, a.kchunk = ((K + 15) / 16) * 16;
  return a;
}

void fill_f32(float* p, int64_t n, float v, hipStream_t stream)
{
  if (n <= 0) return;
  fill_f32_kernel<<<grid_for(static_cast<uint64_t>(n), 256, 1u << 30), 256, 0, stream>>>(p, n, v);
  QPIP_LAUNCH_CHECK();
}

// Forward pass of `rows` (multiple of 16, >= n) rows; fills X, cells, H1, H2, Z.
struct Activations {
  dbuf<uint16_t> X, H1, H2;
  dbuf<uint32_t> cells;
  dbuf<float> Z;
  explicit Activations(int64_t rows)
    : X(rows * kIn), H1(rows * kHid), H2(rows * kHid), cells(rows), Z(rows * kCand)
  {
  }
};

void forward(qpip_ranker const* m, GridView const& g, double const* xy, int64_t n, int64_t rows,
             Activations& act, hipStream_t stream)
{
  Layout const lay(m->num_cells);
  uint16_t const* pb = m->params_bf16;
  {
    KernelTimer kt("ranker_features", stream);
    features_kernel<<<grid_for(static_cast<uint64_t>(rows), 256, 1u << 30), 256, 0, stream>>>(
      g, pb + lay.e, xy, n, rows, act.X.get(), act.cells.get());
    QPIP_LAUNCH_CHECK();
  }
  KernelTimer kt("ranker_forward", stream);
  GemmArgs l1 = gemm(act.X.get(), kIn, 1, pb + lay.w1, 1, kIn, rows, kHid, kIn);
  l1.bias = m->params + lay.b1, l1.relu = 1, l1.Cb = act.H1.get(), l1.cb_m = kHid, l1.cb_n = 1;
  launch_gemm(l1, stream);
  GemmArgs l2 = gemm(act.H1.get(), kHid, 1, pb + lay.w2, 1, kHid, rows, kHid, kHid);
  l2.bias = m->params + lay.b2, l2.relu = 1, l2.Cb = act.H2.get(), l2.cb_m = kHid, l2.cb_n = 1;
  launch_gemm(l2, stream);
  GemmArgs l3 = gemm(act.H2.get(), kHid, 1, pb + lay.w3, 1, kHid, rows, kCand, kHid);
  l3.bias = m->params + lay.b3, l3.C = act.Z.get(), l3.c_m = kCand, l3.c_n = 1;
  launch_gemm(l3, stream);
}

int64_t round16(int64_t n) { return ((n + 15) / 16) * 16; }

}  // namespace

}  // namespace qpip

using namespace qpip;

extern "C" int64_t qpip_ranker_num_params(int64_t num_cells)
{
  return num_cells < 0 ? -1 : Layout(num_cells).total;
}

extern "C" int qpip_ranker_features(const qpip_ranker* model,
                                    const qpip_grid* grid,
                                    const double* xy,
                                    int64_t num_points,
                                    uint16_t* features,
                                    uint32_t* cells,
                                    void* stream_)
{
  return guarded([&] {
    GridView const g = grid_view(grid);
    check_model(model, grid);
    QPIP_EXPECTS(num_points >= 0, "point count must not be negative");
    if (num_points == 0) return;
    auto stream = static_cast<hipStream_t>(stream_);
    StreamScope scope(stream);
    Layout const lay(model->num_cells);
    features_kernel<<<grid_for(static_cast<uint64_t>(num_points), 256, 1u << 30), 256, 0,
                      stream>>>(g, model->params_bf16 + lay.e, xy, num_points, num_points,
                                features, cells);
    QPIP_LAUNCH_CHECK();
  });
}

extern "C" int qpip_ranker_logits(const qpip_ranker* model,
                                  const qpip_grid* grid,
                                  const double* xy,
                                  int64_t num_points,
                                  float* logits,
                                  void* stream_)
{
  return guarded([&] {
    GridView const g = grid_view(grid);
    check_model(model, grid);
    QPIP_EXPECTS(num_points >= 0, "point count must not be negative");
    if (num_points == 0) return;
    auto stream        = static_cast<hipStream_t>(stream_);
    StreamScope scope(stream);
    int64_t const rows = round16(num_points);
    Activations act(rows);
    forward(model, g, xy, num_points, rows, act, stream);
    QPIP_HIP(hipMemcpyAsync(logits, act.Z.get(), num_points * kCand * sizeof(float),
                            hipMemcpyDeviceToDevice, stream));
  });
}

extern "C" int qpip_ranker_gradients(const qpip_ranker* model,
                                     const qpip_grid* grid,
                                     const double* xy,
                                     const int32_t* labels,
                                     int64_t num_points,
                                     float* grad,
                                     float* stats,
                                     void* stream_)
{
  return guarded([&] {
    GridView const g = grid_view(grid);
    check_model(model, grid);
    QPIP_EXPECTS(num_points > 0, "a training batch needs at least one point");
    auto stream        = static_cast<hipStream_t>(stream_);
    StreamScope scope(stream);
    Layout const lay(model->num_cells);
    uint16_t const* pb = model->params_bf16;
    int64_t const rows = round16(num_points);
    Activations act(rows);
    forward(model, g, xy, num_points, rows, act, stream);
    KernelTimer kt("ranker_backward", stream);
    fill_f32(grad, lay.total, 0.f, stream);
    dbuf<uint16_t> dZ(rows * kCand), dH2(rows * kHid), dH1(rows * kHid);
    dbuf<float> dXe(rows * kEmb);
    xent_kernel<<<grid_for(static_cast<uint64_t>(rows), 256, 1u << 30), 256, 0, stream>>>(
      act.Z.get(), act.cells.get(), g.count, labels, num_points, rows,
      1.f / static_cast<float>(num_points), dZ.get(), stats);
    QPIP_LAUNCH_CHECK();
    int64_t const kc = 1024;
    // dW3[o][k] = sum_i dZ[i][o] H2[i][k]
    GemmArgs w3 = gemm(dZ.get(), 1, kCand, act.H2.get(), kHid, 1, kCand, kHid, rows);
    w3.kchunk = kc, w3.atomic = 1, w3.C = grad + lay.w3, w3.c_m = kHid, w3.c_n = 1;
    launch_gemm(w3, stream);
    static_assert(256 % kCand == 0 && 256 % kHid == 0, "colsum_kernel needs cols | 256");
    auto colsum_grid = [&](int cols) {
      return static_cast<unsigned>((rows + colsum_rows_per_block(cols) - 1) /
                                   colsum_rows_per_block(cols));
    };
    colsum_kernel<<<colsum_grid(kCand), 256, 0, stream>>>(dZ.get(), rows, kCand, grad + lay.b3);
    QPIP_LAUNCH_CHECK();
    // dH2 = (dZ W3) * (H2 > 0)
    GemmArgs h2 = gemm(dZ.get(), kCand, 1, pb + lay.w3, kHid, 1, rows, kHid, kCand);
    h2.gate = act.H2.get(), h2.g_m = kHid, h2.g_n = 1, h2.Cb = dH2.get(), h2.cb_m = kHid, h2.cb_n = 1;
    launch_gemm(h2, stream);
    // dW2[o][k] = sum_i dH2[i][o] H1[i][k]
    GemmArgs w2 = gemm(dH2.get(), 1, kHid, act.H1.get(), kHid, 1, kHid, kHid, rows);
    w2.kchunk = kc, w2.atomic = 1, w2.C = grad + lay.w2, w2.c_m = kHid, w2.c_n = 1;
    launch_gemm(w2, stream);
    colsum_kernel<<<colsum_grid(kHid), 256, 0, stream>>>(dH2.get(), rows, kHid, grad + lay.b2);
    QPIP_LAUNCH_CHECK();
    // dH1 = (dH2 W2) * (H1 > 0)
    GemmArgs h1 = gemm(dH2.get(), kHid, 1, pb + lay.w2, kHid, 1, rows, kHid, kHid);
    h1.gate = act.H1.get(), h1.g_m = kHid, h1.g_n = 1, h1.Cb = dH1.get(), h1.cb_m = kHid, h1.cb_n = 1;
    launch_gemm(h1, stream);
    // dW1[o][k] = sum_i dH1[i][o] X[i][k]
    GemmArgs w1 = gemm(dH1.get(), 1, kHid, act.X.get(), kIn, 1, kHid, kIn, rows);
    w1.kchunk = kc, w1.atomic = 1, w1.C = grad + lay.w1, w1.c_m = kIn, w1.c_n = 1;
    launch_gemm(w1, stream);
    colsum_kernel<<<colsum_grid(kHid), 256, 0, stream>>>(dH1.get(), rows, kHid, grad + lay.b1);
    QPIP_LAUNCH_CHECK();
    // dX[:, :64] = dH1 W1[:, :64], scattered into the cells' embeddings
    GemmArgs xe = gemm(dH1.get(), kHid, 1, pb + lay.w1, kIn, 1, rows, kEmb, kHid);
    xe.C = dXe.get(), xe.c_m = kEmb, xe.c_n = 1;
    launch_gemm(xe, stream);
    embed_grad_kernel<<<grid_for(static_cast<uint64_t>(num_points) * kEmb, 256, 1u << 30), 256, 0,
                        stream>>>(dXe.get(), act.cells.get(), num_points, grad + lay.e);
    QPIP_LAUNCH_CHECK();
  });
}

extern "C" int qpip_ranker_adam(float* params,
                                uint16_t* params_bf16,
                                const float* grad,
                                float* m,
                                float* v,
                                int64_t num_params,
                                float lr,
                                float beta1,
                                float beta2,
                                float eps,
                                int64_t step,
                                void* stream_)
{
  return guarded([&] {
    QPIP_EXPECTS(num_params >= 0, "parameter count must not be negative");
    QPIP_EXPECTS(step >= 1, "Adam step counts from 1");
    if (num_params == 0) return;
    auto stream    = static_cast<hipStream_t>(stream_);
    StreamScope scope(stream);
    float const c1 = 1.f / (1.f - std::pow(beta1, static_cast<float>(step)));
    float const c2 = 1.f / (1.f - std::pow(beta2, static_cast<float>(step)));
    KernelTimer kt("ranker_adam", stream);
    adam_kernel<<<grid_for(static_cast<uint64_t>(num_params), 256, 1u << 30), 256, 0, stream>>>(
      params, params_bf16, grad, m, v, num_params, lr, beta1, beta2, eps, c1, c2);
    QPIP_LAUNCH_CHECK();
  });
}

extern "C" int qpip_ranker_locate(const qpip_ranker* model,
                                  const qpip_grid* grid,
                                  const int32_t* poly_offsets,
                                  int64_t num_poly_offsets,
                                  const int32_t* ring_offsets,
                                  int64_t num_ring_offsets,
                                  const double* poly_x,
                                  const double* poly_y,
                                  int64_t num_vertices,
                                  const double* xy,
                                  int64_t num_points,
                                  int32_t max_tests,
                                  int32_t* out_polygon,
                                  uint8_t* out_tests,
                                  uint8_t* out_flags,
                                  void* stream_)
{
  return guarded([&] {
    GridView const g = grid_view(grid);
    check_model(model, grid);
    QPIP_EXPECTS(num_points >= 0, "point count must not be negative");
    QPIP_EXPECTS(max_tests >= 1, "max_tests must be at least 1");
    QPIP_EXPECTS(num_poly_offsets >= 1 && num_ring_offsets >= 1,
                 "offset columns need at least one entry");
    if (num_points == 0) return;
    auto stream           = static_cast<hipStream_t>(stream_);
    StreamScope scope(stream);
    int64_t const polys   = num_poly_offsets - 1;
    int64_t const rings   = num_ring_offsets - 1;
    dbuf<Edge> edges(std::max<int64_t>(num_vertices, 1));
    dbuf<EdgeB> edges_b(std::max<int64_t>(num_vertices, 1));
    dbuf<uint32_t> ring_edges(std::max<int64_t>(rings, 1));
    if (rings > 0) {
      KernelTimer kt("ranker_edges", stream);
      build_edges_kernel<<<grid_for(static_cast<uint64_t>(rings), 256 / kWave), 256, 0, stream>>>(
        reinterpret_cast<uint32_t const*>(ring_offsets), rings, poly_x, poly_y, num_vertices,
        edges.get(), edges_b.get(), ring_edges.get(), nullptr, nullptr, nullptr);
      QPIP_LAUNCH_CHECK();
    }
    LocateArgs a{};
    a.g   = g;
    a.pb  = model->params_bf16;
    a.p   = model->params;
    a.lay = Layout(model->num_cells);
    a.P   = PolyView{reinterpret_cast<uint32_t const*>(poly_offsets), polys,
                   reinterpret_cast<uint32_t const*>(ring_offsets), rings, edges.get(),
                   edges_b.get(), ring_edges.get()};
    a.xy        = xy;
    a.n         = num_points;
    a.max_tests = max_tests;
    a.out_poly  = out_polygon;
    a.out_tests = out_tests;
    a.out_flags = out_flags;
    dbuf<Record> records(std::max<int64_t>(polys, 1));
    a.P.records = records.get();
    if (polys > 0) {
      records_kernel<<<grid_for(static_cast<uint64_t>(polys), 256), 256, 0, stream>>>(
        a.P, records.get());
      QPIP_LAUNCH_CHECK();
    }
    dbuf<uint16_t> order(num_points);
    a.order = order.get();
    int dev = 0, cus = 0;
    QPIP_HIP(hipGetDevice(&dev));
    QPIP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    uint64_t const groups = (static_cast<uint64_t>(num_points) + kWave - 1) / kWave;
    uint64_t const want   = (groups + kLocWaves - 1) / kLocWaves;
    unsigned const blocks = static_cast<unsigned>(std::min<uint64_t>(want, uint64_t(2) * cus));
    {
      KernelTimer kt("ranker_rank", stream);
      rank_kernel<<<blocks, kLocWaves * kWave, 0, stream>>>(a);
      QPIP_LAUNCH_CHECK();
    }
    {
      KernelTimer kt("ranker_verify", stream);
      verify_kernel<<<grid_for(static_cast<uint64_t>(num_points), 256, 1u << 30), 256, 0, stream>>>(a);
      QPIP_LAUNCH_CHECK();
    }
  });
}
