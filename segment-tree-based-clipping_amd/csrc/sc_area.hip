// sc_area.hip -- differentiable intersection area and IoU of a batch of jobs
// (row A19; the reference has no counterpart: it only writes the rings).
//
// Forward, per job: A_I = sum of the unsigned shoelace areas of the job's
// cleaned output rings, A_P / A_Q = unsigned shoelace areas of its base /
// overlay ring, IoU = A_I / (A_P + A_Q - A_I).  Twice the signed area of a
// ring is S = sum_i (x_i y_{i+1} - x_{i+1} y_i) and its area is |S| / 2: the
// input rings may run either way round, and with a clockwise input one job's
// output can hold rings of both orientations.  An intersection of two simple
// rings has no holes, so every output ring counts positively.  Each sum is
// reduced in a fixed order, so repeated runs give the same bits.
//
// Backward, with the combinatorics of the forward held fixed: an output
// vertex gets d|S|/dx_i = sgn(S) (y_{i+1} - y_{i-1}), d|S|/dy_i =
// sgn(S) (x_{i-1} - x_{i+1}) with S its ring's signed sum, routed by its
// provenance (the augmented slot it was emitted from):
//  - a source vertex of either ring, or an intersection located at an input
//    vertex (types 2-8: at b1 or c1 as the intersection stage places them),
//    passes the gradient to that input vertex;
//  - a proper crossing (type 1) X = (1-alpha) b1 + alpha b2 with
//    alpha = AP1 / (AP1 - AP2), AP1 = A(b1,c1,c2), AP2 = A(b2,c1,c2)
//    (lineSegmentIntersection2, oacc_segment_tree.cpp:5520-5611) passes it to
//    b1, b2 directly and through alpha to b1, b2, c1, c2.
// Contributions meet at input vertices through FP64 atomic adds.
#include "sc_ctx.hpp"
#include "sc_fp64.hpp"

#include <climits>

namespace sc {

namespace {

constexpr int BLK = 256;
constexpr int SRC_NONE = INT_MIN;

inline unsigned grid_for(int64_t n, int per = BLK) {
    int64_t g = ceil_div(n, per);
    if (g < 1) g = 1;
    return unsigned(g);
}

__device__ inline double block_sum(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int d = BLK / 2; d > 0; d >>= 1) {
        if (int(threadIdx.x) < d) sh[threadIdx.x] += sh[threadIdx.x + d];
        __syncthreads();
    }
    return sh[0];
}

// twice the signed area of input ring r (vertices voff[r] .. voff[r+1]-1, the
// last one the repeated closing vertex); one block per ring
__global__ __launch_bounds__(BLK) void k_ring_area2(const double *__restrict__ poly, const int *__restrict__ voff,
                                                    int r0, double *out) {
    __shared__ double sh[BLK];
    const int r = r0 + blockIdx.x;
    const int64_t a = voff[r], n = voff[r + 1] - a - 1;
    // thread t adds the terms of vertices t, t + BLK, ... in that order; four
    // terms are loaded before they are added (the same order and bits)
    double s = 0.0;
    int64_t i = threadIdx.x;
    for (; i + 3 * BLK < n; i += 4 * BLK) {
        const double *p0 = poly + 2 * (a + i), *p1 = p0 + 2 * BLK, *p2 = p1 + 2 * BLK, *p3 = p2 + 2 * BLK;
        const double t0 = p0[0] * p0[3] - p0[2] * p0[1], t1 = p1[0] * p1[3] - p1[2] * p1[1];
        const double t2 = p2[0] * p2[3] - p2[2] * p2[1], t3 = p3[0] * p3[3] - p3[2] * p3[1];
        s += t0;
        s += t1;
        s += t2;
        s += t3;
    }
    for (; i < n; i += BLK) {
        const double *p = poly + 2 * (a + i);
        s += p[0] * p[3] - p[2] * p[1];
    }
    double t = block_sum(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// Twice the signed area of every output ring.  A ring of at most AREA_CHUNK
// vertices is summed by one wave: lanes stride the ring and a fixed shuffle
// tree adds their sums (the order of the small-pair kernel); the waves of a
// fixed grid take rings in turn, so a batch of many small rings costs no
// launch of a wave per ring.  A longer ring is summed in chunks of
// AREA_CHUNK vertices the same way, one wave per chunk, and its chunk sums
// are added in chunk order.
constexpr int AREA_CHUNK = 64 * WAVE;
constexpr int AREA_WAVES = 8192; // waves of the short-ring grid (32 per CU)

__device__ inline double wave_ring_sum(const double *__restrict__ x, const double *__restrict__ y, int64_t a,
                                       int64_t b, int64_t c0, int64_t c1, int lane) {
    double s = 0.0;
    for (int64_t i = c0 + lane; i < c1; i += WAVE) {
        const int64_t nx = i + 1 == b ? a : i + 1;
        s += x[i] * y[nx] - x[nx] * y[i];
    }
    for (int d = WAVE / 2; d > 0; d >>= 1) s += __shfl_down(s, d, WAVE);
    return s;
}

__global__ __launch_bounds__(BLK) void k_out_ring_area2(int nrings, const int *__restrict__ roff,
                                                        const double *__restrict__ x, const double *__restrict__ y,
                                                        double *out) {
    const int lane = int(threadIdx.x) & (WAVE - 1);
    const int stride = int(gridDim.x) * (BLK / WAVE);
    for (int r = int(blockIdx.x) * (BLK / WAVE) + int(threadIdx.x) / WAVE; r < nrings; r += stride) {
        const int64_t a = roff[r], b = roff[r + 1];
        if (b - a > AREA_CHUNK) continue; // chunked below
        const double s = wave_ring_sum(x, y, a, b, a, b, lane);
        if (lane == 0) out[r] = s;
    }
}

// chunks of each long ring (0 for a ring of at most AREA_CHUNK vertices;
// cnt[nrings] = 0 closes the scan)
__global__ void k_ring_chunks(int nrings, const int *__restrict__ roff, uint32_t *cnt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nrings) return;
    if (r == nrings) {
        cnt[r] = 0u;
        return;
    }
    const int64_t len = roff[r + 1] - roff[r];
    cnt[r] = len <= AREA_CHUNK ? 0u : uint32_t((len + AREA_CHUNK - 1) / AREA_CHUNK);
}

// chunk q of the *total (choff: exclusive scan of the chunk counts)
__global__ __launch_bounds__(BLK) void k_out_chunk_area2(int nrings, const int *__restrict__ roff,
                                                         const uint32_t *__restrict__ choff,
                                                         const uint32_t *__restrict__ total, int64_t cap,
                                                         const double *__restrict__ x, const double *__restrict__ y,
                                                         double *part) {
    const int64_t q = int64_t(blockIdx.x) * (BLK / WAVE) + int64_t(threadIdx.x) / WAVE;
    const int lane = int(threadIdx.x) & (WAVE - 1);
    if (q >= int64_t(*total) || q >= cap) return; // whole wave
    int lo = 0, hi = nrings - 1; // the ring: the first r with choff[r + 1] > q
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (int64_t(choff[mid + 1]) > q) hi = mid;
        else lo = mid + 1;
    }
    const int64_t a = roff[lo], b = roff[lo + 1];
    const int64_t c0 = a + (q - int64_t(choff[lo])) * AREA_CHUNK;
    const int64_t c1 = c0 + AREA_CHUNK < b ? c0 + AREA_CHUNK : b;
    const double s = wave_ring_sum(x, y, a, b, c0, c1, lane);
    if (lane == 0) part[q] = s;
}

__global__ void k_ring_chunk_sum(int nrings, const uint32_t *__restrict__ choff, int64_t cap,
                                 const double *__restrict__ part, double *out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrings) return;
    const int64_t k0 = choff[r], k1 = int64_t(choff[r + 1]) < cap ? int64_t(choff[r + 1]) : cap;
    if (k0 >= k1) return; // a short ring
    double s = part[k0];
    for (int64_t k = k0 + 1; k < k1; ++k) s += part[k];
    out[r] = s;
}

// twice the unsigned area of each job's output: one block per job, thread t
// adds |S| of rings t, t + 256, ... of the job, then a fixed tree
__global__ __launch_bounds__(BLK) void k_job_out_area2(const int *__restrict__ jring,
                                                       const double *__restrict__ ring2, double *out) {
    __shared__ double sh[BLK];
    const int j = blockIdx.x;
    const int r0 = jring[j], r1 = jring[j + 1];
    double s = 0.0;
    for (int r = r0 + int(threadIdx.x); r < r1; r += BLK) s += fabs(ring2[r]);
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) out[j] = t;
}

__global__ void k_job_final(int J, int paired, const double *out2, const double *barea2, const double *carea2,
                            double *out4) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= J) return;
    double I = 0.5 * out2[j], P = 0.5 * fabs(barea2[paired ? j : 0]), Q = 0.5 * fabs(carea2[j]);
    double U = P + Q - I;
    out4[j] = I;
    out4[J + j] = P;
    out4[2 * J + j] = Q;
    out4[3 * J + j] = U != 0.0 ? I / U : 0.0;
}

__device__ inline double sgn1(double v) { return v < 0.0 ? -1.0 : 1.0; }

// per job: coefficients of d(2A_I) (per output ring, times its orientation
// sign in k_output_grad), d(S_P) and d(S_Q) in the loss, S the signed sums
__global__ void k_coef(int J, int paired, const double *out4, const double *barea2, const double *carea2,
                       const double *g_area, const double *g_iou, double *cI, double *cQ, double *cPr) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= J) return;
    double I = out4[j], P = out4[J + j], Q = out4[2 * J + j];
    double U = P + Q - I;
    double ga = g_area ? g_area[j] : 0.0, gi = g_iou ? g_iou[j] : 0.0;
    double dI = ga, dP = 0.0;
    if (U != 0.0) {
        dI += gi * (U + I) / (U * U);
        dP = -gi * I / (U * U);
    }
    cI[j] = 0.5 * dI;
    cQ[j] = 0.5 * dP * sgn1(carea2[j]);
    // the base ring may be shared by every job (one-to-many)
    unsafeAtomicAdd(&cPr[paired ? j : 0], 0.5 * dP * sgn1(barea2[paired ? j : 0]));
}

// d(2A)/d(vertex) of the input rings: base rings (all of them, coefficient per
// ring) or overlays (coefficient per overlay, zero outside the batch); input
// vertex i of ring r sits at i + r in the layout with closing vertices.
// Plain stores: this is the first writer of the gradient arrays.
__global__ void k_input_ring_grad(int64_t n, const int *__restrict__ off, int R, const double *__restrict__ poly,
                                  const double *__restrict__ coef, double *grad) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = seg_of(off, R, i);
    const int64_t v = i - off[r], m = off[r + 1] - off[r];
    const double *P = poly + 2 * (int64_t(off[r]) + r);
    const int64_t nx = v + 1 == m ? 0 : v + 1, pv = v == 0 ? m - 1 : v - 1;
    const double c = coef[r];
    grad[2 * i] = c * (P[2 * nx + 1] - P[2 * pv + 1]);
    grad[2 * i + 1] = c * (P[2 * pv] - P[2 * nx]);
}

struct ProvArgs {
    int J, nrings;
    const int *jring, *roff, *src;
    const uint32_t *rid; // output ring of every output vertex
    const double *ox, *oy;
    const double *cI;
    const double *ring2; // signed 2A of every output ring
    const int *sv_b, *sp_b, *sv_c, *sp_c;
    const int *p_b, *p_c, *p_type;
    const int *bvoff, *beoff, *nb, *cvoff, *cin, *nc; // per job
    const double *bpoly, *cpoly;
    double *grad_b, *grad_c;
    int *bad;
};

__device__ inline void add_b(const ProvArgs &a, int j, int v, double gx, double gy) {
    int64_t i = int64_t(a.beoff[j]) + v;
    unsafeAtomicAdd(&a.grad_b[2 * i], gx);
    unsafeAtomicAdd(&a.grad_b[2 * i + 1], gy);
}
__device__ inline void add_c(const ProvArgs &a, int j, int v, double gx, double gy) {
    int64_t i = int64_t(a.cin[j]) + v;
    unsafeAtomicAdd(&a.grad_c[2 * i], gx);
    unsafeAtomicAdd(&a.grad_c[2 * i + 1], gy);
}

// gradient (gx, gy) on the point of pair p, seen from side P (onP) or Q
__device__ void pair_grad(const ProvArgs &a, int j, int p, bool onP, double gx, double gy) {
    const int t = a.p_type[p], b = a.p_b[p], c = a.p_c[p];
    if (t != 1) {
        // types 3,5,7 sit at c1 (base side), 2,6 at b1 (overlay side)
        if (onP) add_c(a, j, c, gx, gy);
        else add_b(a, j, b, gx, gy);
        return;
    }
    const int nb = a.nb[j], nc = a.nc[j];
    const double *B = a.bpoly + 2 * (int64_t(a.bvoff[j]) + b);
    const double *C = a.cpoly + 2 * (int64_t(a.cvoff[j]) + c);
    const double b1x = B[0], b1y = B[1], b2x = B[2], b2y = B[3];
    const double c1x = C[0], c1y = C[1], c2x = C[2], c2y = C[3];
    const double AP1 = areaA(b1x, b1y, c1x, c1y, c2x, c2y);
    const double AP2 = areaA(b2x, b2y, c1x, c1y, c2x, c2y);
    const double D = AP1 - AP2;
    const double al = AP1 / D;
    // X = (1-al) b1 + al b2
    const double gal = gx * (b2x - b1x) + gy * (b2y - b1y);
    double gb1x = (1.0 - al) * gx, gb1y = (1.0 - al) * gy, gb2x = al * gx, gb2y = al * gy;
    // al = AP1 / (AP1 - AP2)
    const double g1 = -gal * AP2 / (D * D), g2 = gal * AP1 / (D * D);
    // A(p,q,r): d/dp = (qy-ry, rx-qx), d/dq = (ry-py, px-rx), d/dr = (py-qy, qx-px)
    gb1x += g1 * (c1y - c2y);
    gb1y += g1 * (c2x - c1x);
    gb2x += g2 * (c1y - c2y);
    gb2y += g2 * (c2x - c1x);
    const double gc1x = g1 * (c2y - b1y) + g2 * (c2y - b2y);
    const double gc1y = g1 * (b1x - c2x) + g2 * (b2x - c2x);
    const double gc2x = g1 * (b1y - c1y) + g2 * (b2y - c1y);
    const double gc2y = g1 * (c1x - b1x) + g2 * (c1x - b2x);
    add_b(a, j, b, gb1x, gb1y);
    add_b(a, j, b + 1 == nb ? 0 : b + 1, gb2x, gb2y);
    add_c(a, j, c, gc1x, gc1y);
    add_c(a, j, c + 1 == nc ? 0 : c + 1, gc2x, gc2y);
}

// ring heads for the output ring index (rid zero-filled, then max-scanned)
__global__ void k_out_ring_heads(int nrings, const int *__restrict__ roff, uint32_t *rid) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrings) return;
    if (roff[r + 1] > roff[r]) rid[roff[r]] = uint32_t(r);
}

__global__ void k_output_grad(ProvArgs a, int64_t v0, int64_t v1) {
    int64_t i = v0 + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= v1) return;
    const int r = int(a.rid[i]);
    const int j = seg_of(a.jring, a.J, r);
    const int64_t rb = a.roff[r], re = a.roff[r + 1];
    const int64_t nx = i + 1 == re ? rb : i + 1, pv = i == rb ? re - 1 : i - 1;
    const double c = a.cI[j] * sgn1(a.ring2[r]);
    const double gx = c * (a.oy[nx] - a.oy[pv]), gy = c * (a.ox[pv] - a.ox[nx]);
    const int s = a.src[i];
    if (s == SRC_NONE) {
        atomicOr(a.bad, 1);
        return;
    }
    if (s >= 0) {
        const int v = a.sv_b[s];
        if (v >= 0) add_b(a, j, v, gx, gy);
        else if (a.sp_b[s] >= 0) pair_grad(a, j, a.sp_b[s], true, gx, gy);
        else atomicOr(a.bad, 1);
    } else {
        const int g = ~s;
        const int v = a.sv_c[g];
        if (v >= 0) add_c(a, j, v, gx, gy);
        else if (a.sp_c[g] >= 0) pair_grad(a, j, a.sp_c[g], false, gx, gy);
        else atomicOr(a.bad, 1);
    }
}

struct AreaScratch {
    DBuf<double> out2, ring2, barea2, carea2, out4, cI, cQ, cPr, part;
    DBuf<int> bad, bin_off;
    DBuf<uint32_t> chunks, choff, tot, rid;
};

AreaScratch &ascratch(Ctx &) {
    thread_local AreaScratch *p = new AreaScratch();
    return *p;
}

void job_rings_to_device(Ctx &c) {
    JobSet &js = c.jobs;
    SC_REQUIRE(js.J >= 1 && int(js.ring_off.size()) == js.J + 1, SC_ESTATE, "no clipped batch");
    js.d_ring.ensure(js.J + 1);
    SC_HIP(hipMemcpyAsync(js.d_ring.p, js.ring_off.data(), sizeof(int) * (js.J + 1), hipMemcpyHostToDevice,
                          c.stream));
}

} // namespace

void batch_area(Ctx &c, double *out4) {
    SC_REQUIRE(!c.uni, SC_ESTATE, "area / IoU are defined for the intersection (a union output may hold holes)");
    JobSet &js = c.jobs;
    job_rings_to_device(c);
    hipStream_t s = c.stream;
    AreaScratch &a = ascratch(c);
    const int J = js.J;
    double *out2 = a.out2.ensure(J), *barea2 = a.barea2.ensure(c.nbr), *carea2 = a.carea2.ensure(J);
    double *ring2 = a.ring2.ensure(size_t(c.nrings > 0 ? c.nrings : 1));
    if (c.nrings > 0) {
        const int nr = c.nrings;
        const int64_t wgrid = std::min<int64_t>(ceil_div(int64_t(nr), BLK / WAVE), AREA_WAVES / (BLK / WAVE));
        k_out_ring_area2<<<unsigned(wgrid), BLK, 0, s>>>(nr, c.d_ring_off, c.d_ring_x, c.d_ring_y, ring2);
        uint32_t *cnt = a.chunks.ensure(size_t(nr) + 1), *choff = a.choff.ensure(size_t(nr) + 1);
        uint32_t *tot = a.tot.ensure(1);
        // a long ring has at most one chunk more than its AREA_CHUNK-vertex
        // multiples, and there are at most n_out / AREA_CHUNK long rings
        const int64_t cap = 2 * (c.n_out / AREA_CHUNK) + 1;
        double *part = a.part.ensure(size_t(cap));
        k_ring_chunks<<<grid_for(int64_t(nr) + 1), BLK, 0, s>>>(nr, c.d_ring_off, cnt);
        scan_exclusive_u32(cnt, choff, int64_t(nr) + 1, tot, s);
        k_out_chunk_area2<<<grid_for(cap, BLK / WAVE), BLK, 0, s>>>(nr, c.d_ring_off, choff, tot, cap, c.d_ring_x,
                                                                    c.d_ring_y, part);
        k_ring_chunk_sum<<<grid_for(nr), BLK, 0, s>>>(nr, choff, cap, part, ring2);
    }
    k_job_out_area2<<<J, BLK, 0, s>>>(js.d_ring, ring2, out2);
    if (c.paired) k_ring_area2<<<J, BLK, 0, s>>>(c.bpoly, c.d_bvoff, js.k0 - 1, barea2);
    else k_ring_area2<<<1, BLK, 0, s>>>(c.bpoly, c.d_bvoff, 0, barea2);
    k_ring_area2<<<J, BLK, 0, s>>>(c.cpoly, c.d_coff, js.k0 - 1, carea2);
    k_job_final<<<grid_for(J), BLK, 0, s>>>(J, c.paired ? 1 : 0, out2, barea2, carea2, out4);
    SC_HIP(hipGetLastError());
}

void batch_area_backward(Ctx &c, const double *g_area, const double *g_iou, double *grad_b, double *grad_c) {
    JobSet &js = c.jobs;
    SC_REQUIRE(c.prov, SC_ESTATE, "the batch was clipped without provenance (segclip_set_provenance)");
    hipStream_t s = c.stream;
    AreaScratch &a = ascratch(c);
    const int J = js.J;
    double *out4 = a.out4.ensure(4 * size_t(J));
    batch_area(c, out4);
    double *cI = a.cI.ensure(J), *cQ = a.cQ.ensure(c.K), *cPr = a.cPr.ensure(c.nbr);
    SC_HIP(hipMemsetAsync(cPr, 0, sizeof(double) * c.nbr, s));
    SC_HIP(hipMemsetAsync(cQ, 0, sizeof(double) * c.K, s));
    // k_coef writes cQ for the batch's jobs at their overlay index
    k_coef<<<grid_for(J), BLK, 0, s>>>(J, c.paired ? 1 : 0, out4, a.barea2, a.carea2, g_area, g_iou, cI,
                                       cQ + (js.k0 - 1), c.paired ? cPr + (js.k0 - 1) : cPr);
    SC_HIP(hipGetLastError());
    // input ring terms: base rings (input offsets beoff), all overlays (input offsets eoff - nb)
    int *boff = c.d_beoff;
    k_input_ring_grad<<<grid_for(c.nb), BLK, 0, s>>>(c.nb, boff, c.nbr, c.bpoly, cPr, grad_b);
    std::vector<int> cinall(size_t(c.K) + 1);
    for (int k = 0; k <= c.K; ++k) cinall[k] = c.eoff[k] - c.nb;
    int *dci = a.bin_off.ensure(c.K + 1);
    SC_HIP(hipMemcpyAsync(dci, cinall.data(), sizeof(int) * (c.K + 1), hipMemcpyHostToDevice, s));
    k_input_ring_grad<<<grid_for(c.nct), BLK, 0, s>>>(c.nct, dci, c.K, c.cpoly, cQ, grad_c);
    SC_HIP(hipGetLastError());
    // output vertices of the batch
    ProvArgs p;
    p.J = J;
    p.nrings = c.nrings;
    p.jring = js.d_ring;
    p.roff = c.d_ring_off;
    p.src = c.d_ring_src;
    p.ox = c.d_ring_x;
    p.oy = c.d_ring_y;
    p.cI = cI;
    p.ring2 = a.ring2;
    p.sv_b = c.sv_b;
    p.sp_b = c.sp_b;
    p.sv_c = c.sv_c;
    p.sp_c = c.sp_c;
    p.p_b = c.p_b;
    p.p_c = c.p_c;
    p.p_type = c.p_type;
    p.bvoff = js.d_bvoff;
    p.beoff = js.d_beoff;
    p.nb = js.d_nb;
    p.cvoff = js.d_cvoff;
    p.cin = js.d_cin;
    p.nc = js.d_nc;
    p.bpoly = c.bpoly;
    p.cpoly = c.cpoly;
    p.grad_b = grad_b;
    p.grad_c = grad_c;
    p.bad = a.bad.ensure(1);
    SC_HIP(hipMemsetAsync(p.bad, 0, sizeof(int), s));
    if (c.n_out > 0 && c.nrings > 0) {
        // the ring of every output vertex: ring heads, then a max-scan
        uint32_t *rid = a.rid.ensure(size_t(c.n_out));
        {
            FillBatch fb(s);
            fb.add(rid, c.n_out, 0u);
            fb.flush();
        }
        k_out_ring_heads<<<grid_for(c.nrings), BLK, 0, s>>>(c.nrings, c.d_ring_off, rid);
        scan_inclusive_max_u32(rid, rid, c.n_out, s);
        p.rid = rid;
        k_output_grad<<<grid_for(c.n_out), BLK, 0, s>>>(p, 0, c.n_out);
    }
    SC_HIP(hipGetLastError());
    int bad = 0;
    {
        ReadBack rb(c);
        rb.add(&bad, static_cast<const int *>(p.bad), 1);
        rb.run();
    }
    SC_REQUIRE(!bad, SC_EUNDEF, "an output vertex has no provenance (sequential route midpoint vertex)");
}

} // namespace sc
