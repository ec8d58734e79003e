// Differentiable form of the step (include/pbf.h, pbf_adjoint_*): reverse
// mode through predict, the I x (lambda, position update) iterations, update,
// XSPH and vorticity confinement, with each step's sort order and neighbour
// table held fixed (computed by the plain step's kernels from the state).
//
// The step is differentiated as the race-free solver computes it:
//   x*   = x + dt (v + dt g gdir)                      predictpos.glsl:18-38
//   p0   = x*[perm]                                    sorted by cell key
//   lam  = -(rho/rho0 - 1) / (sum |g_j|^2 + |sum g_j|^2 + eps), g_j = gradW/rho0
//   p'   = clamp(p + 1/rho0 sum (lam_i + lam_j + s_ij) gradW_ij)   (I times)
//   v    = (p_I - x) / dt ;  x' = p_I                  update.glsl:16-28
//   vx   = v + c sum (v_j - v_i) W_ij ; w = sum (v_j - v_i) x gradW_ij
//   eta  = sum |w_j| gradW_ij ; N = eta / |eta| ; v' = vx + dt eps_v (N x w)
// with W = poly6, gradW the spiky gradient, s_ij = -k (scale W)^4. The
// derivative of the clamp is 1 inside the walls and 0 where it clamps; N has a
// zero derivative where eta = 0; neighbours past n (a zero record, reached
// only through a wrapped count) are constants.
//
// Reverse mode visits each particle's neighbours and scatters the neighbour
// side of every pair term with atomic adds, so it needs no symmetry of the
// neighbour sets (the reference's are not always symmetric). Element type T is
// float (the production gradient) or double (the finite-difference check).
// Forward states are checkpointed every K steps; the backward pass recomputes
// each segment's steps from its checkpoint into a tape, then runs the segment
// in reverse. The last segment is taped by the forward pass itself.
#include <cstring>
#include <type_traits>
#include <new>
#include <vector>

#include "pbf_internal.h"

namespace pbf {
namespace {

constexpr int AB = 256;  // workgroup size of the adjoint kernels

template <class T>
struct alignas(sizeof(T) * 4 > 32 ? 32 : sizeof(T) * 4) T4 {
    T x, y, z, w;
};

// Forward-mode dual number in double: a value and its derivative along one
// direction of the inputs. The forward kernels run on it unchanged, which
// gives directional derivatives of the FP64 map without any finite step
// (pbf_adjoint_jvp): an AD path independent of the hand-written reverse
// kernels. Comparisons and the clamp act on the value, so the derivative is
// that of the same smooth piece the reverse pass differentiates.
struct Dual {
    double v, d;
    __host__ __device__ Dual() : v(0.0), d(0.0) {}
    __host__ __device__ Dual(double x) : v(x), d(0.0) {}
    __host__ __device__ Dual(int x) : v((double)x), d(0.0) {}
    __host__ __device__ Dual(double x, double dx) : v(x), d(dx) {}
    __host__ __device__ explicit operator float() const { return (float)v; }
    __device__ Dual &operator+=(const Dual &b) {
        v += b.v;
        d += b.d;
        return *this;
    }
    __device__ Dual &operator-=(const Dual &b) {
        v -= b.v;
        d -= b.d;
        return *this;
    }
    __device__ Dual &operator*=(const Dual &b) {
        d = d * b.v + v * b.d;
        v *= b.v;
        return *this;
    }
};
__device__ __forceinline__ Dual operator+(Dual a, const Dual &b) { return a += b; }
__device__ __forceinline__ Dual operator-(Dual a, const Dual &b) { return a -= b; }
__device__ __forceinline__ Dual operator-(const Dual &a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, const Dual &b) { return a *= b; }
__device__ __forceinline__ Dual operator/(const Dual &a, const Dual &b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ bool operator<(const Dual &a, const Dual &b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(const Dual &a, const Dual &b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<=(const Dual &a, const Dual &b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>=(const Dual &a, const Dual &b) { return a.v >= b.v; }
__device__ __forceinline__ bool operator==(const Dual &a, const Dual &b) { return a.v == b.v; }
__device__ __forceinline__ bool operator!=(const Dual &a, const Dual &b) { return a.v != b.v; }
// (the overloads of the element types stay visible next to the dual one)
__device__ __forceinline__ float sqrt(float x) { return ::sqrtf(x); }
__device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
__device__ __forceinline__ Dual sqrt(const Dual &a) {
    const double s = ::sqrt(a.v);
    return Dual(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
}

template <class T>
__device__ __forceinline__ T sqlen(T x, T y, T z) {
    return (z * z + y * y) + x * x;
}

// Per-step constants in T.
template <class T>
struct AConsts {
    int32_t n, gx, gy, gz;
    T h, hh, h6, h9, inv_rho0, eps, gravity, dt, k, scale, xsph_c, vort_eps;
    T lo[3], hi[3], gdir[3];
    int32_t extforce;
    int32_t vorticity;
};

// 315/(64 pi) and -3 * 15/pi (calclambda.glsl:46, 63): the shaders' FP32
// constants for float, the same decimal constants in double for the FP64 form.
template <class T>
__device__ __forceinline__ T poly6c();
template <>
__device__ __forceinline__ float poly6c<float>() { return 1.56668147106f; }
template <>
__device__ __forceinline__ double poly6c<double>() { return 1.56668147106; }
template <class T>
__device__ __forceinline__ T spikyc();
template <>
__device__ __forceinline__ float spikyc<float>() { return -3.0f * 4.774648292756860f; }
template <>
__device__ __forceinline__ double spikyc<double>() { return -3.0 * 4.774648292756860; }
template <>
__device__ __forceinline__ Dual poly6c<Dual>() { return Dual(1.56668147106); }
template <>
__device__ __forceinline__ Dual spikyc<Dual>() { return Dual(-3.0 * 4.774648292756860); }

// W_poly6 of sq = r^2 (r <= h).
template <class T>
__device__ __forceinline__ T wpoly(const AConsts<T> &c, T sq) {
    const T t = c.hh - sq;
    return (t * t) * (t * poly6c<T>()) / c.h9;
}
// dW/dd = -6 poly6 (hh - sq)^2 / h^9 * d  (scalar factor)
template <class T>
__device__ __forceinline__ T dwpoly(const AConsts<T> &c, T sq) {
    const T t = c.hh - sq;
    return (T)-6.0 * poly6c<T>() * t * t / c.h9;
}
// spiky gradient factor: gradW = sg * d, sg = spiky (h - r)^2 / (r h^6)
template <class T>
__device__ __forceinline__ T gspiky(const AConsts<T> &c, T r) {
    const T t = c.h - r;
    return ((spikyc<T>() * t) * t) / (r * c.h6);
}
// gradW (calclambda.glsl:57-64) rounded as the plain step computes it:
// ((spiky t) t) d / (r h h h h h h), so the float forward equals the plain step.
template <class T>
__device__ __forceinline__ void grad3(const AConsts<T> &c, T r, T ex, T ey, T ez, T &gx, T &gy, T &gz) {
    const T t = c.h - r;
    const T s = (spikyc<T>() * t) * t;
    const T den = (((((r * c.h) * c.h) * c.h) * c.h) * c.h) * c.h;
    gx = (s * ex) / den;
    gy = (s * ey) / den;
    gz = (s * ez) / den;
}
// Jacobian of gradW: J a = sp/h^6 [f a + (f'/r) (d.a) d], f = (h-r)^2/r,
// f'/r = -(h-r)(h+r)/r^3.
template <class T>
__device__ __forceinline__ void jspiky(const AConsts<T> &c, T r, T dx, T dy, T dz, T ax, T ay, T az, T &ox, T &oy,
                                       T &oz) {
    const T t = c.h - r;
    const T f = t * t / r;
    const T fr = -t * (c.h + r) / (r * r * r);
    const T s = spikyc<T>() / c.h6;
    const T da = dx * ax + dy * ay + dz * az;
    ox = s * (f * ax + fr * da * dx);
    oy = s * (f * ay + fr * da * dy);
    oz = s * (f * az + fr * da * dz);
}

template <class T>
__device__ __forceinline__ T4<T> load_nb(const T4<T> *p, int32_t j, int32_t n) {
    return j < n ? p[j] : T4<T>{0, 0, 0, 0};
}

template <class T>
__device__ __forceinline__ void atomic_add3(T4<T> *p, int32_t j, T x, T y, T z) {
    atomicAdd(&p[j].x, x);
    atomicAdd(&p[j].y, y);
    atomicAdd(&p[j].z, z);
}

// FOR_EACH_NEIGHBOUR (foreachneighbour.glsl:1-10) over the fixed table:
// body(j, p[j]) for every candidate j != i of the 9 rows, in the reference's
// order, with the positions of eight candidates loaded together (indices past
// n read the zero record). At the adjoint's sizes a launch has few waves, each
// walking its candidates serially, so load latency rather than arithmetic
// sets the time.
constexpr int kAdjBatch = 8;

template <class T, class F>
__device__ __forceinline__ void each_candidate(const int32_t *__restrict__ nbr9, int32_t n, int32_t i,
                                               const T4<T> *__restrict__ p, F &&body) {
    for (int o = 0; o < 9; o++) {
        const int32_t d = nbr9[(size_t)o * n + i];
        const int32_t s0 = d & 0xFFFFFF, e = d >> 24;
        const int32_t last = max(min(s0 + e, n) - 1, 0);
        for (int32_t q = 0; q < e; q += kAdjBatch) {
            T4<T> pj[kAdjBatch];
#pragma unroll
            for (int u = 0; u < kAdjBatch; u++) {
                const int32_t j = s0 + q + u;
                pj[u] = p[min(j, last)];
                if (j >= n) pj[u] = T4<T>{0, 0, 0, 0};
            }
#pragma unroll
            for (int u = 0; u < kAdjBatch; u++) {
                const int32_t j = s0 + q + u;
                if (q + u < e && j != i) body(j, pj[u]);
            }
        }
    }
}

// The reverse kernels run a group of kAdjLanes consecutive lanes of one wave
// per particle; lane sub of the group takes the particle's pair terms sub,
// sub + kAdjLanes, ... of the recorded pair list (the neighbours within h of
// the forward, in list order), or without one (none recorded, or more than
// the list holds) the candidates sub, sub + kAdjLanes, ... of each row. The
// group's sums are then reduced across its lanes (group_sum), so a particle's
// own accumulators take one atomic add per launch. extra(j) loads the
// neighbour's operands of the kernel with the batch's positions, so
// kAdjBatch neighbours of a lane are in flight together.
constexpr int kAdjLanes = 8;
static_assert(64 % kAdjLanes == 0 && AB % kAdjLanes == 0, "lane groups inside a wave");

template <class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int m = 1; m < kAdjLanes; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

template <class T, class X, class E, class F>
__device__ __forceinline__ void each_pair(const uint16_t *__restrict__ ents, const uint16_t *__restrict__ cnts,
                                          const int32_t *__restrict__ nbr9, int32_t n, int32_t i, int32_t sub,
                                          const T4<T> *__restrict__ p, E &&extra, F &&body) {
    const uint32_t cnt = cnts ? cnts[i] : kListNone;
    if (cnt == kListNone) {
        for (int o = 0; o < 9; o++) {
            const int32_t d = nbr9[(size_t)o * n + i];
            const int32_t s0 = d & 0xFFFFFF, e = d >> 24;
            for (int32_t q = sub; q < e; q += kAdjLanes) {
                const int32_t j = s0 + q;
                if (j != i) body(j, load_nb(p, j, n), extra(j));
            }
        }
        return;
    }
    for (uint32_t e0 = (uint32_t)sub; e0 < cnt; e0 += kAdjBatch * kAdjLanes) {
        int32_t j[kAdjBatch];
        T4<T> pj[kAdjBatch];
        X xj[kAdjBatch];
#pragma unroll
        for (int u = 0; u < kAdjBatch; u++) {
            const uint32_t e = e0 + (uint32_t)(u * kAdjLanes);
            const uint32_t code = ents[(size_t)(e < cnt ? e : e0) * n + i];
            j[u] = (nbr9[(size_t)(code >> 8) * n + i] & 0xFFFFFF) + (int32_t)(code & 255u);
        }
#pragma unroll
        for (int u = 0; u < kAdjBatch; u++) {
            pj[u] = load_nb(p, j[u], n);
            xj[u] = extra(j[u]);
        }
#pragma unroll
        for (int u = 0; u < kAdjBatch; u++)
            if (e0 + (uint32_t)(u * kAdjLanes) < cnt) body(j[u], pj[u], xj[u]);
    }
}

// Pair terms without atomics. The neighbour side of a pair term (what lane i
// would add to particle j) is, for every pair of these kernels, the same
// expression as j's own side of the reversed pair with the operands swapped:
// when j lists i whenever i lists j, lane i can gather the neighbour sides
// meant for it from its own list instead of receiving them by atomic adds.
// That holds for two particles whose nine rows are exactly the union of their
// 27 cells' contents (proper, ka_proper): the range test is symmetric (the
// squared distance has the same bits both ways) and cell adjacency is. A pair
// with an improper particle (the reference's quirk cells, wrapped counts,
// rows at the grid's faces) keeps the scattered form in both directions.

__device__ __forceinline__ void block_add2(double a, double b, double *acc) {
    __shared__ double s[2][AB];
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    __syncthreads();
    for (int k = AB / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            s[0][threadIdx.x] += s[0][threadIdx.x + k];
            s[1][threadIdx.x] += s[1][threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (s[0][0] != 0.0) atomicAdd(&acc[0], s[0][0]);
        if (s[1][0] != 0.0) atomicAdd(&acc[1], s[1][0]);
    }
}

// ------------------------------------------------------------------ forward
// predictpos.glsl:18-38 in T; also the float record {x*, id} and its key for
// the structure (sort and cell table) of this step.
template <class T>
__global__ __launch_bounds__(AB) void ka_predict(AConsts<T> c, const T4<T> *__restrict__ x, const T4<T> *__restrict__ v,
                                                 T4<T> *__restrict__ xs, float4 *__restrict__ xs_f) {
    const int32_t i = blockIdx.x * AB + threadIdx.x;
    if (i >= c.n) return;
    const T4<T> p = x[i], u = v[i];
    T vx = u.x, vy = u.y, vz = u.z;
    if (c.extforce && p.z > (T)c.gz / (T)2) {
        // velocity += 2 gravity (0, 0, -1) dt on every component
        // (predictpos.glsl:26-27): the zero terms turn a -0 velocity into +0
        const T g2 = (T)2 * c.gravity;
        vx = vx + (g2 * (T)0) * c.dt;
        vy = vy + (g2 * (T)0) * c.dt;
        vz = vz + (g2 * (T)-1) * c.dt;
    }
    vx = vx + (c.gravity * c.gdir[0]) * c.dt;
    vy = vy + (c.gravity * c.gdir[1]) * c.dt;
    vz = vz + (c.gravity * c.gdir[2]) * c.dt;
    const T a = p.x + c.dt * vx, b = p.y + c.dt * vy, d = p.z + c.dt * vz;
    xs[i] = T4<T>{a, b, d, 0};
    xs_f[i] = make_float4((float)a, (float)b, (float)d, __int_as_float(i));
}

// True cell ranges of the sorted keys, without the reference's table quirks:
// tstart[c] and tend[c] of every occupied cell c (index x + gx (y + gy z),
// decoded from the key x + gx z + gx gz y). Keys are exact cell hashes here
// (the grid has fewer than 2^24 cells); tstart is -1 elsewhere.
__global__ __launch_bounds__(AB) void ka_true_cells(int32_t n, int32_t gx, int32_t gy, int32_t gz,
                                                    const uint32_t *__restrict__ skeys, int32_t *__restrict__ tstart,
                                                    int32_t *__restrict__ tend) {
    const int32_t g = blockIdx.x * AB + threadIdx.x;
    if (g >= n) return;
    const uint32_t k = skeys[g];
    const uint32_t x = k % (uint32_t)gx, t = k / (uint32_t)gx, z = t % (uint32_t)gz, y = t / (uint32_t)gz;
    if (y >= (uint32_t)gy) return;
    const size_t c = (size_t)x + (size_t)gx * ((size_t)y + (size_t)gy * z);
    if (g == 0 || skeys[g - 1] != k) tstart[c] = g;
    if (g == n - 1 || skeys[g + 1] != k) tend[c] = g + 1;
}

// proper[i]: particle i (sorted slot, record rec[i]) lies in the cell its key
// names, its 27 cells are inside the grid, and each of its nine row words is
// the one the true cell ranges give (first start, total count below 128), so
// its candidates are exactly the particles of its 27 cells.
__global__ __launch_bounds__(AB) void ka_proper(int32_t n, int32_t gx, int32_t gy, int32_t gz, int32_t exact,
                                                const float4 *__restrict__ rec, const uint32_t *__restrict__ skeys,
                                                const int32_t *__restrict__ tstart, const int32_t *__restrict__ tend,
                                                const int32_t *__restrict__ nbr9, uint8_t *__restrict__ proper) {
    const int32_t i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    const float4 r = rec[i];
    bool ok = exact != 0 && r.x >= 1.0f && r.y >= 1.0f && r.z >= 1.0f && r.x < (float)(gx - 1) &&
              r.y < (float)(gy - 1) && r.z < (float)(gz - 1);
    if (ok) {
        const int32_t px = (int32_t)r.x, py = (int32_t)r.y, pz = (int32_t)r.z;  // as the neighbour-cell pass
        ok = skeys[i] == (uint32_t)px + (uint32_t)gx * ((uint32_t)pz + (uint32_t)gz * (uint32_t)py);
        // the 27 cells' words in two rounds of loads in flight together
        int32_t w[9], s0[27], e0[27];
#pragma unroll
        for (int o = 0; o < 9; o++) w[o] = nbr9[(size_t)o * n + i];
#pragma unroll
        for (int q = 0; q < 27; q++) {
            const int32_t y = py + q / 9 - 1, z = pz + (q / 3) % 3 - 1, x = px + q % 3 - 1;
            s0[q] = tstart[(size_t)x + (size_t)gx * ((size_t)y + (size_t)gy * (size_t)z)];
        }
#pragma unroll
        for (int q = 0; q < 27; q++) {
            const int32_t y = py + q / 9 - 1, z = pz + (q / 3) % 3 - 1, x = px + q % 3 - 1;
            e0[q] = s0[q] == -1 ? 0 : tend[(size_t)x + (size_t)gx * ((size_t)y + (size_t)gy * (size_t)z)];
        }
#pragma unroll
        for (int o = 0; o < 9; o++) {  // row o: (dy, dz) = (o / 3 - 1, o % 3 - 1), cells q = 3 o + dx + 1
            int32_t first = -1, cnt = 0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int q = 3 * o + k;
                if (s0[q] == -1) continue;
                if (first == -1) first = s0[q];
                cnt += e0[q] - s0[q];
            }
            ok = ok && cnt < 128 && w[o] == (int32_t)((uint32_t)first + ((uint32_t)cnt << 24));
        }
    }
    proper[i] = ok ? 1 : 0;
}

template <class T>
__global__ __launch_bounds__(AB) void ka_gather(int32_t n, const uint32_t *__restrict__ perm,
                                                const T4<T> *__restrict__ xs, T4<T> *__restrict__ p) {
    const int32_t g = blockIdx.x * AB + threadIdx.x;
    if (g >= n) return;
    p[g] = xs[perm[g]];
}

// calclambda.glsl:66-103; keeps S + eps and G = sum g_j for the reverse pass.
template <class T>
__global__ __launch_bounds__(AB) void ka_lambda(AConsts<T> c, const T4<T> *__restrict__ p,
                                                const int32_t *__restrict__ nbr9, T *__restrict__ lam,
                                                T *__restrict__ sden_out, T4<T> *__restrict__ G_out) {
    const int32_t i = blockIdx.x * AB + threadIdx.x;
    if (i >= c.n) return;
    const T4<T> pi = p[i];
    T rho = 0, sk = 0, gx = 0, gy = 0, gz = 0;
    each_candidate(nbr9, c.n, i, p, [&](int32_t j, const T4<T> &pj) {
        const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const T sq = sqlen(dx, dy, dz);
        const T r = sqrt(sq);
        if (r > c.h) return;
        rho += wpoly(c, sq);
        if (r == (T)0) return;
        T ax, ay, az;
        grad3(c, r, dx, dy, dz, ax, ay, az);
        ax = ax * c.inv_rho0;
        ay = ay * c.inv_rho0;
        az = az * c.inv_rho0;
        sk += sqlen(ax, ay, az);
        gx += ax;
        gy += ay;
        gz += az;
    });
    sk += sqlen(gx, gy, gz);
    const T ci = rho * c.inv_rho0 - (T)1;
    lam[i] = -ci / (sk + c.eps);
    if (sden_out) {
        sden_out[i] = sk + c.eps;
        G_out[i] = T4<T>{gx, gy, gz, 0};
    }
}

// updatepos.glsl:42-105, race-free. Clamp state per particle (mask): bit a
// set where axis a was not clamped, bit 3 + a where it was clamped to the
// upper wall. frozen (a forward's recorded states, or null): every axis takes
// the recorded decision instead of its own (the map whose derivative the
// reverse pass computes), and flips counts the decisions the clamp would have
// made otherwise.
template <class T>
__global__ __launch_bounds__(AB) void ka_updatepos(AConsts<T> c, const T4<T> *__restrict__ p,
                                                   const T *__restrict__ lam, const int32_t *__restrict__ nbr9,
                                                   T4<T> *__restrict__ out, uint8_t *__restrict__ mask,
                                                   const uint8_t *__restrict__ frozen,
                                                   unsigned long long *__restrict__ flips) {
    const int32_t i = blockIdx.x * AB + threadIdx.x;
    if (i >= c.n) return;
    const T4<T> pi = p[i];
    const T li = lam[i];
    T dx = 0, dy = 0, dz = 0;
    each_candidate(nbr9, c.n, i, p, [&](int32_t j, const T4<T> &pj) {
        const T ex = pi.x - pj.x, ey = pi.y - pj.y, ez = pi.z - pj.z;
        const T sq = sqlen(ex, ey, ez);
        const T r = sqrt(sq);
        if (r > c.h || r == (T)0) return;
        T sc = c.scale * wpoly(c, sq);
        sc *= sc;
        sc *= sc;
        sc = -c.k * sc;
        const T lj = j < c.n ? lam[j] : (T)0;
        T gx, gy, gz;
        grad3(c, r, ex, ey, ez, gx, gy, gz);
        const T f = li + lj + sc;
        dx += f * gx;
        dy += f * gy;
        dz += f * gz;
    });
    const T q[3] = {pi.x + c.inv_rho0 * dx, pi.y + c.inv_rho0 * dy, pi.z + c.inv_rho0 * dz};
    T o[3];
    uint8_t m = 0;
    for (int a = 0; a < 3; a++) {
        o[a] = q[a] < c.lo[a] ? c.lo[a] : (q[a] > c.hi[a] ? c.hi[a] : q[a]);
        if (q[a] >= c.lo[a] && q[a] <= c.hi[a]) m |= (uint8_t)(1u << a);
        else if (q[a] > c.hi[a]) m |= (uint8_t)(8u << a);
    }
    if (frozen) {
        const uint8_t f = frozen[i];
        unsigned k = 0;  // axes whose decision (inside, lower wall, upper wall) differs
        for (int a = 0; a < 3; a++) {
            const unsigned df = ((f >> a) & 1u) ? 0u : (((f >> (3 + a)) & 1u) ? 2u : 1u);
            const unsigned dm = ((m >> a) & 1u) ? 0u : (((m >> (3 + a)) & 1u) ? 2u : 1u);
            o[a] = df == 0u ? q[a] : (df == 2u ? c.hi[a] : c.lo[a]);
            k += df != dm;
        }
        if (k) atomicAdd(flips, (unsigned long long)k);
        m = f;
    }
    out[i] = T4<T>{o[0], o[1], o[2], pi.w};
    if (mask) mask[i] = m;
}

// update.glsl:16-28: v = (p_I - x)/dt by id; x = p_I; vs keeps v in sorted order.
template <class T>
__global__ __launch_bounds__(AB) void ka_update(AConsts<T> c, const uint32_t *__restrict__ perm,
                                                const T4<T> *__restrict__ p, T4<T> *__restrict__ x,
                                                T4<T> *__restrict__ v, T4<T> *__restrict__ vs) {
    const int32_t g = blockIdx.x * AB + threadIdx.x;
    if (g >= c.n) return;
    const uint32_t id = perm[g];
    const T4<T> r = p[g], old = x[id];
    const T4<T> u = T4<T>{(r.x - old.x) / c.dt, (r.y - old.y) / c.dt, (r.z - old.z) / c.dt, 0};
    x[id] = T4<T>{r.x, r.y, r.z, 0};
    v[id] = u;
    vs[g] = u;
}

// vorticity.glsl:34-60: XSPH velocity and omega; |omega| per particle.
template <class T>
__global__ __launch_bounds__(AB) void ka_vort_a(AConsts<T> c, const T4<T> *__restrict__ p,
                                                const int32_t *__restrict__ nbr9, const T4<T> *__restrict__ vs,
                                                const T4<T> *__restrict__ v_id0, T4<T> *__restrict__ vx,
                                                T4<T> *__restrict__ om, T *__restrict__ o) {
    const int32_t i = blockIdx.x * AB + threadIdx.x;
    if (i >= c.n) return;
    const T4<T> pi = p[i], vi = vs[i];
    T sx = 0, sy = 0, sz = 0, wx = 0, wy = 0, wz = 0;
    each_candidate(nbr9, c.n, i, p, [&](int32_t j, const T4<T> &pj) {
        const T ex = pi.x - pj.x, ey = pi.y - pj.y, ez = pi.z - pj.z;
        const T sq = sqlen(ex, ey, ez);
        const T r = sqrt(sq);
        if (r > c.h) return;
        const T4<T> vj = j < c.n ? vs[j] : v_id0[0];
        const T ux = vj.x - vi.x, uy = vj.y - vi.y, uz = vj.z - vi.z;
        const T w = wpoly(c, sq);
        sx += ux * w;
        sy += uy * w;
        sz += uz * w;
        if (r == (T)0) return;
        T gx, gy, gz;
        grad3(c, r, ex, ey, ez, gx, gy, gz);
        wx = (wx + uy * gz) - uz * gy;
        wy = (wy + uz * gx) - ux * gz;
        wz = (wz + ux * gy) - uy * gx;
    });
    vx[i] = T4<T>{vi.x + c.xsph_c * sx, vi.y + c.xsph_c * sy, vi.z + c.xsph_c * sz, 0};
    om[i] = T4<T>{wx, wy, wz, 0};
    o[i] = sqrt(sqlen(wx, wy, wz));
}

// vorticity.glsl:65-85: eta, N and the confinement; v by id.
template <class T>
__global__ __launch_bounds__(AB) void ka_vort_b(AConsts<T> c, const uint32_t *__restrict__ perm,
                                                const T4<T> *__restrict__ p, const int32_t *__restrict__ nbr9,
                                                const T4<T> *__restrict__ vx, const T4<T> *__restrict__ om,
                                                const T *__restrict__ o, T4<T> *__restrict__ v,
                                                T4<T> *__restrict__ eta_out) {
    const int32_t i = blockIdx.x * AB + threadIdx.x;
    if (i >= c.n) return;
    const T4<T> pi = p[i];
    T nx = 0, ny = 0, nz = 0;
    each_candidate(nbr9, c.n, i, p, [&](int32_t j, const T4<T> &pj) {
        const T ex = pi.x - pj.x, ey = pi.y - pj.y, ez = pi.z - pj.z;
        const T sq = sqlen(ex, ey, ez);
        const T r = sqrt(sq);
        if (r > c.h || r == (T)0) return;
        const T oj = j < c.n ? o[j] : (T)0;
        T gx, gy, gz;
        grad3(c, r, ex, ey, ez, gx, gy, gz);
        nx += oj * gx;
        ny += oj * gy;
        nz += oj * gz;
    });
    if (eta_out) eta_out[i] = T4<T>{nx, ny, nz, 0};
    const T l = sqrt(sqlen(nx, ny, nz));
    if (l > (T)0) {
        nx = nx / l;
        ny = ny / l;
        nz = nz / l;
    }
    const T4<T> w = om[i], u = vx[i];
    const T sc = c.dt * c.vort_eps;
    v[perm[i]] = T4<T>{u.x + sc * (ny * w.z - nz * w.y), u.y + sc * (nz * w.x - nx * w.z),
                       u.z + sc * (nx * w.y - ny * w.x), 0};
}

// ------------------------------------------------------------------ reverse
// vorticity pass B reversed, per particle: from vbar (by id) to vxbar, the
// own part of ombar, d/d eps_v and etabar = d/d eta (eta from the forward's
// tape), for the pair pass that follows.
template <class T>
__global__ __launch_bounds__(AB) void kr_vort_b_own(AConsts<T> c, const uint32_t *__restrict__ perm,
                                                    const T4<T> *__restrict__ om, const T4<T> *__restrict__ eta,
                                                    const T4<T> *__restrict__ vbar, T4<T> *__restrict__ vxbar,
                                                    T4<T> *__restrict__ ombar, T4<T> *__restrict__ etabar,
                                                    double *__restrict__ acc) {
    const int32_t i = blockIdx.x * AB + threadIdx.x;
    double deps = 0.0;
    if (i < c.n) {
        const T4<T> et = eta[i];
        const T l = sqrt(sqlen(et.x, et.y, et.z));
        T Nx = 0, Ny = 0, Nz = 0;
        if (l > (T)0) {
            Nx = et.x / l;
            Ny = et.y / l;
            Nz = et.z / l;
        }
        const T4<T> w = om[i];
        const T4<T> a = vbar[perm[i]];
        const T sc = c.dt * c.vort_eps;
        vxbar[i] = T4<T>{a.x, a.y, a.z, 0};
        // d/d eps_v of a . (dt eps (N x w))
        const T cx = Ny * w.z - Nz * w.y, cy = Nz * w.x - Nx * w.z, cz = Nx * w.y - Ny * w.x;
        deps = (double)(c.dt * (a.x * cx + a.y * cy + a.z * cz));
        // a . (N x w) = w . (a x N)
        ombar[i] = T4<T>{sc * (a.y * Nz - a.z * Ny), sc * (a.z * Nx - a.x * Nz), sc * (a.x * Ny - a.y * Nx), 0};
        // a . (N x w) = N . (w x a); N = eta / |eta|: etabar = (Nbar - N (N . Nbar)) / |eta|
        T4<T> eb{0, 0, 0, 0};
        if (l > (T)0) {
            const T Nbx = sc * (w.y * a.z - w.z * a.y), Nby = sc * (w.z * a.x - w.x * a.z),
                    Nbz = sc * (w.x * a.y - w.y * a.x);
            const T nn = Nx * Nbx + Ny * Nby + Nz * Nbz;
            eb = T4<T>{(Nbx - Nx * nn) / l, (Nby - Ny * nn) / l, (Nbz - Nz * nn) / l, 0};
        }
        etabar[i] = eb;
    }
    block_add2(0.0, deps, acc);
}

template <class T>
struct VortBX {
    T4<T> eb;
    T o;
    uint8_t proper;
};

// eta_i = sum_j o_j gradW_ij reversed: obar and pbar. Own side of pair (i, j):
// obar_j += s (d . eb_i), pbar_i += o_j J eb_i, pbar_j -= o_j J eb_i (d = p_i - p_j,
// J the Jacobian of gradW, even in d). Gathered from (j, i): obar_i -= s (d . eb_j),
// pbar_i -= o_i J eb_j.
template <class T>
__global__ __launch_bounds__(AB) void kr_vort_b(AConsts<T> c, const T4<T> *__restrict__ p,
                                                const int32_t *__restrict__ nbr9, const uint16_t *__restrict__ ents,
                                                const uint16_t *__restrict__ cnts, const uint8_t *__restrict__ proper,
                                                const T *__restrict__ o, const T4<T> *__restrict__ etabar,
                                                T *__restrict__ obar, T4<T> *__restrict__ pbar) {
    const int64_t tid = (int64_t)blockIdx.x * AB + threadIdx.x;
    const int32_t i = (int32_t)(tid / kAdjLanes), sub = (int32_t)(tid % kAdjLanes);
    if (i >= c.n) return;  // whole groups
    const T4<T> pi = p[i];
    const T4<T> eb = etabar[i];
    const T oi = o[i];
    const bool pri = proper[i] != 0;
    T ob = 0, pbx = 0, pby = 0, pbz = 0;
    each_pair<T, VortBX<T>>(
        ents, cnts, nbr9, c.n, i, sub, p,
        [&](int32_t j) {
            return j < c.n ? VortBX<T>{etabar[j], o[j], proper[j]} : VortBX<T>{T4<T>{0, 0, 0, 0}, (T)0, (uint8_t)0};
        },
        [&](int32_t j, const T4<T> &pj, const VortBX<T> &xj) {
            const T ex = pi.x - pj.x, ey = pi.y - pj.y, ez = pi.z - pj.z;
            const T sq = sqlen(ex, ey, ez);
            const T r = sqrt(sq);
            if (r > c.h || r == (T)0) return;
            const T s = gspiky(c, r);
            T jx, jy, jz;
            if (pri && xj.proper) {
                ob -= s * (ex * xj.eb.x + ey * xj.eb.y + ez * xj.eb.z);
                jspiky(c, r, ex, ey, ez, xj.o * eb.x - oi * xj.eb.x, xj.o * eb.y - oi * xj.eb.y,
                       xj.o * eb.z - oi * xj.eb.z, jx, jy, jz);
                pbx += jx;
                pby += jy;
                pbz += jz;
                return;
            }
            if (j < c.n) atomicAdd(&obar[j], s * (ex * eb.x + ey * eb.y + ez * eb.z));
            jspiky(c, r, ex, ey, ez, eb.x, eb.y, eb.z, jx, jy, jz);
            pbx += xj.o * jx;
            pby += xj.o * jy;
            pbz += xj.o * jz;
            if (j < c.n) atomic_add3(pbar, j, -xj.o * jx, -xj.o * jy, -xj.o * jz);
        });
    ob = group_sum(ob);
    pbx = group_sum(pbx);
    pby = group_sum(pby);
    pbz = group_sum(pbz);
    if (sub != 0) return;
    if (ob != (T)0) atomicAdd(&obar[i], ob);
    atomic_add3(pbar, i, pbx, pby, pbz);
}

// |omega| reversed: ombar += obar omega / |omega|.
template <class T>
__global__ __launch_bounds__(AB) void kr_omega_norm(int32_t n, const T4<T> *__restrict__ om, const T *__restrict__ o,
                                                    const T *__restrict__ obar, T4<T> *__restrict__ ombar) {
    const int32_t i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    const T oi = o[i];
    if (oi > (T)0) {
        const T f = obar[i] / oi;
        const T4<T> w = om[i];
        T4<T> b = ombar[i];
        b.x += f * w.x;
        b.y += f * w.y;
        b.z += f * w.z;
        ombar[i] = b;
    }
}

template <class T>
struct VortAX {
    T4<T> v, a, wb;
    uint8_t proper;
};

// vorticity pass A reversed: from vxbar and ombar to vsbar and pbar, and
// d/d xsph_c. Own side of pair (i, j), u = v_j - v_i, g = gradW_ij:
// ub = c W a_i + g x wb_i goes to vsbar_j (and -ub to vsbar_i); the position
// term d = c dW/dd (u . a_i) + J (wb_i x u) to pbar_i (and -d to pbar_j).
// Gathered from (j, i) (u, d and g change sign, W and J do not): vsbar_i gets
// c W a_j - g x wb_j, pbar_i gets -(c dW/dd (u . a_j) - J (wb_j x u)).
template <class T>
__global__ __launch_bounds__(AB) void kr_vort_a(AConsts<T> c, const T4<T> *__restrict__ p,
                                                const int32_t *__restrict__ nbr9, const uint16_t *__restrict__ ents,
                                                const uint16_t *__restrict__ cnts, const uint8_t *__restrict__ proper,
                                                const T4<T> *__restrict__ vs, const T4<T> *__restrict__ vxbar,
                                                const T4<T> *__restrict__ ombar, T4<T> *__restrict__ vsbar,
                                                T4<T> *__restrict__ pbar, double *__restrict__ acc) {
    const int64_t tid = (int64_t)blockIdx.x * AB + threadIdx.x;
    const int32_t i = (int32_t)(tid / kAdjLanes), sub = (int32_t)(tid % kAdjLanes);
    double dc = 0.0;
    if (i < c.n) {  // whole groups
        const T4<T> pi = p[i], vi = vs[i], a = vxbar[i], wb = ombar[i];
        const bool pri = proper[i] != 0;
        T sx = 0, sy = 0, sz = 0;
        T vbx = sub == 0 ? a.x : (T)0, vby = sub == 0 ? a.y : (T)0, vbz = sub == 0 ? a.z : (T)0;  // vx = v + ...
        T pbx = 0, pby = 0, pbz = 0;
        each_pair<T, VortAX<T>>(
            ents, cnts, nbr9, c.n, i, sub, p,
            [&](int32_t j) {
                const T4<T> z{0, 0, 0, 0};
                return j < c.n ? VortAX<T>{vs[j], vxbar[j], ombar[j], proper[j]} : VortAX<T>{z, z, z, (uint8_t)0};
            },
            [&](int32_t j, const T4<T> &pj, const VortAX<T> &xj) {
                const T ex = pi.x - pj.x, ey = pi.y - pj.y, ez = pi.z - pj.z;
                const T sq = sqlen(ex, ey, ez);
                const T r = sqrt(sq);
                if (r > c.h) return;
                if (j >= c.n) return;  // zero record: its velocity is a constant here
                const T ux = xj.v.x - vi.x, uy = xj.v.y - vi.y, uz = xj.v.z - vi.z;
                const T W = wpoly(c, sq);
                sx += ux * W;
                sy += uy * W;
                sz += uz * W;
                const bool sym = pri && xj.proper;
                // the a and wb of the terms: own a_i, wb_i; with the gathered side a_i - a_j, wb_i + wb_j
                const T dax = sym ? a.x - xj.a.x : a.x, day = sym ? a.y - xj.a.y : a.y, daz = sym ? a.z - xj.a.z : a.z;
                const T wx = sym ? wb.x + xj.wb.x : wb.x, wy = sym ? wb.y + xj.wb.y : wb.y,
                        wz = sym ? wb.z + xj.wb.z : wb.z;
                // ub = c W a + g x wb, for vsbar (own side negated on i)
                T ubx = c.xsph_c * W * dax, uby = c.xsph_c * W * day, ubz = c.xsph_c * W * daz;
                const T dw = dwpoly(c, sq) * c.xsph_c * (ux * dax + uy * day + uz * daz);
                T dx = dw * ex, dy = dw * ey, dz = dw * ez;
                if (r > (T)0) {
                    const T s = gspiky(c, r);
                    const T gx = s * ex, gy = s * ey, gz = s * ez;
                    ubx += gy * wz - gz * wy;
                    uby += gz * wx - gx * wz;
                    ubz += gx * wy - gy * wx;
                    // wb . (u x gradW) = gradW . (wb x u)
                    const T qx = wy * uz - wz * uy, qy = wz * ux - wx * uz, qz = wx * uy - wy * ux;
                    T jx, jy, jz;
                    jspiky(c, r, ex, ey, ez, qx, qy, qz, jx, jy, jz);
                    dx += jx;
                    dy += jy;
                    dz += jz;
                }
                pbx += dx;
                pby += dy;
                pbz += dz;
                if (sym) {
                    // -(c W a_i + g x wb_i) + (c W a_j - g x wb_j)
                    vbx -= ubx;
                    vby -= uby;
                    vbz -= ubz;
                    return;
                }
                atomic_add3(vsbar, j, ubx, uby, ubz);
                vbx -= ubx;
                vby -= uby;
                vbz -= ubz;
                atomic_add3(pbar, j, -dx, -dy, -dz);
            });
        dc = (double)(a.x * sx + a.y * sy + a.z * sz);  // this lane's share of a . sum u W
        vbx = group_sum(vbx);
        vby = group_sum(vby);
        vbz = group_sum(vbz);
        pbx = group_sum(pbx);
        pby = group_sum(pby);
        pbz = group_sum(pbz);
        if (sub == 0) {
            atomic_add3(vsbar, i, vbx, vby, vbz);
            atomic_add3(pbar, i, pbx, pby, pbz);
        }
    }
    block_add2(dc, 0.0, acc);
}


// update reversed: pbar_I += xbar[id] + vsbar/dt ; xbar_old[id] = -vsbar/dt.
// Without vorticity vsbar is the incoming vbar[id].
template <class T>
__global__ __launch_bounds__(AB) void kr_update(AConsts<T> c, const uint32_t *__restrict__ perm,
                                                const T4<T> *__restrict__ xbar_new, const T4<T> *__restrict__ vbar_new,
                                                const T4<T> *__restrict__ vsbar, T4<T> *__restrict__ pbar,
                                                T4<T> *__restrict__ xbar_old) {
    const int32_t g = blockIdx.x * AB + threadIdx.x;
    if (g >= c.n) return;
    const uint32_t id = perm[g];
    const T4<T> vb = vsbar ? vsbar[g] : vbar_new[id];
    const T4<T> xb = xbar_new[id];
    T4<T> pb = pbar[g];
    pb.x += xb.x + vb.x / c.dt;
    pb.y += xb.y + vb.y / c.dt;
    pb.z += xb.z + vb.z / c.dt;
    pbar[g] = pb;
    xbar_old[id] = T4<T>{-vb.x / c.dt, -vb.y / c.dt, -vb.z / c.dt, 0};
}

template <class T>
struct UpdX {
    T4<T> b;
    T lam;
    uint8_t mask, proper;
};

// position update reversed: pbar_prev = mask pbar_next + pair terms; lambar.
// Own side of pair (i, j), a_i = the clamp-masked pbar_next of i:
// gdot = (1/rho0) gradW . a_i to lambar_i and lambar_j; d = (1/rho0) f J a_i +
// gdot ds/dd to pbar_i (-d to pbar_j), f = lam_i + lam_j + s_corr (symmetric).
// The gathered side of (j, i) is the same with -a_j: the terms are linear in a.
template <class T>
__global__ __launch_bounds__(AB) void kr_updatepos(AConsts<T> c, const T4<T> *__restrict__ p,
                                                   const T *__restrict__ lam, const int32_t *__restrict__ nbr9,
                                                   const uint16_t *__restrict__ ents,
                                                   const uint16_t *__restrict__ cnts,
                                                   const uint8_t *__restrict__ proper,
                                                   const uint8_t *__restrict__ mask, const T4<T> *__restrict__ pbn,
                                                   T4<T> *__restrict__ pbp, T *__restrict__ lambar) {
    const int64_t tid = (int64_t)blockIdx.x * AB + threadIdx.x;
    const int32_t i = (int32_t)(tid / kAdjLanes), sub = (int32_t)(tid % kAdjLanes);
    if (i >= c.n) return;  // whole groups
    const T4<T> pi = p[i];
    const uint8_t m = mask[i];
    const T4<T> b = pbn[i];
    const T ax = (m & 1) ? b.x : (T)0, ay = (m & 2) ? b.y : (T)0, az = (m & 4) ? b.z : (T)0;
    const bool pri = proper[i] != 0;
    T pbx = sub == 0 ? ax : (T)0, pby = sub == 0 ? ay : (T)0, pbz = sub == 0 ? az : (T)0;  // the identity part
    const T li = lam[i];
    T lb = 0;
    const T k4 = (T)-4 * c.k * c.scale * c.scale * c.scale * c.scale;
    each_pair<T, UpdX<T>>(
        ents, cnts, nbr9, c.n, i, sub, p,
        [&](int32_t j) {
            return j < c.n ? UpdX<T>{pbn[j], lam[j], mask[j], proper[j]}
                           : UpdX<T>{T4<T>{0, 0, 0, 0}, (T)0, (uint8_t)0, (uint8_t)0};
        },
        [&](int32_t j, const T4<T> &pj, const UpdX<T> &xj) {
            const T ex = pi.x - pj.x, ey = pi.y - pj.y, ez = pi.z - pj.z;
            const T sq = sqlen(ex, ey, ez);
            const T r = sqrt(sq);
            if (r > c.h || r == (T)0) return;
            const T W = wpoly(c, sq);
            T sc = c.scale * W;
            sc *= sc;
            sc *= sc;
            sc = -c.k * sc;
            const T f = li + xj.lam + sc;
            const T s = gspiky(c, r);
            const bool sym = pri && xj.proper;
            // a_i, or a_i - a_j with the gathered side
            T vx = ax, vy = ay, vz = az;
            if (sym) {
                vx -= (xj.mask & 1) ? xj.b.x : (T)0;
                vy -= (xj.mask & 2) ? xj.b.y : (T)0;
                vz -= (xj.mask & 4) ? xj.b.z : (T)0;
            }
            const T gdot = c.inv_rho0 * s * (ex * vx + ey * vy + ez * vz);  // (1/rho0) gradW . v
            lb += gdot;
            if (!sym && j < c.n) atomicAdd(&lambar[j], gdot);
            // d/dd of (1/rho0) f gradW . v: f J v + (gradW . v) ds/dd
            T jx, jy, jz;
            jspiky(c, r, ex, ey, ez, vx, vy, vz, jx, jy, jz);
            const T ds = k4 * W * W * W * dwpoly(c, sq);  // ds/dd = -4k scale^4 W^3 dW/dd (factor of d)
            const T dx = c.inv_rho0 * f * jx + gdot * ds * ex;
            const T dy = c.inv_rho0 * f * jy + gdot * ds * ey;
            const T dz = c.inv_rho0 * f * jz + gdot * ds * ez;
            pbx += dx;
            pby += dy;
            pbz += dz;
            if (!sym && j < c.n) atomic_add3(pbp, j, -dx, -dy, -dz);
        });
    lb = group_sum(lb);
    pbx = group_sum(pbx);
    pby = group_sum(pby);
    pbz = group_sum(pbz);
    if (sub != 0) return;
    if (lb != (T)0) atomicAdd(&lambar[i], lb);
    atomic_add3(pbp, i, pbx, pby, pbz);
}

template <class T>
struct LamX {
    T4<T> G;
    T lb, S, lam;
    uint8_t proper;
};

// lambda reversed: from lambar to pbar. lam = -C/S: rhobar = -lb/(rho0 S),
// Sbar = -lam lb / S. Own side of pair (i, j): d = rb_i dW/dd + (1/rho0) J gbar_i,
// gbar_i = 2 Sbar_i (g_ij + G_i) (g = gradW / rho0), to pbar_i (-d to pbar_j).
// The gathered side of (j, i) has g_ji = -g_ij and dW/dd odd in d: with it,
// pbar_i gets (rb_i + rb_j) dW/dd + (1/rho0) J (2 (Sbar_i + Sbar_j) g_ij
// + 2 (Sbar_i G_i - Sbar_j G_j)).
template <class T>
__global__ __launch_bounds__(AB) void kr_lambda(AConsts<T> c, const T4<T> *__restrict__ p,
                                                const int32_t *__restrict__ nbr9, const uint16_t *__restrict__ ents,
                                                const uint16_t *__restrict__ cnts,
                                                const uint8_t *__restrict__ proper, const T *__restrict__ lam,
                                                const T *__restrict__ sden, const T4<T> *__restrict__ G,
                                                const T *__restrict__ lambar, T4<T> *__restrict__ pbar,
                                                T4<T> *__restrict__ clear_p, T *__restrict__ clear_l) {
    const int64_t tid = (int64_t)blockIdx.x * AB + threadIdx.x;
    const int32_t i = (int32_t)(tid / kAdjLanes), sub = (int32_t)(tid % kAdjLanes);
    if (i >= c.n) return;  // whole groups
    // the next iteration's accumulators: the position update's input of this
    // iteration (read by its launch, which has finished) and the other lambar
    if (sub == 0) {
        clear_p[i] = T4<T>{0, 0, 0, 0};
        clear_l[i] = (T)0;
    }
    const T lb = lambar[i];
    const bool pri = proper[i] != 0;
    // with no gathered side, a particle without lambar adds nothing (the
    // whole group leaves)
    if (lb == (T)0 && !pri) return;
    const T4<T> pi = p[i];
    const T S = sden[i];
    const T rb = -lb * c.inv_rho0 / S;
    const T sb = -lam[i] * lb / S;
    const T4<T> Gi = G[i];
    T pbx = 0, pby = 0, pbz = 0;
    each_pair<T, LamX<T>>(
        ents, cnts, nbr9, c.n, i, sub, p,
        [&](int32_t j) {
            return j < c.n ? LamX<T>{G[j], lambar[j], sden[j], lam[j], proper[j]}
                           : LamX<T>{T4<T>{0, 0, 0, 0}, (T)0, (T)1, (T)0, (uint8_t)0};
        },
        [&](int32_t j, const T4<T> &pj, const LamX<T> &xj) {
            const T ex = pi.x - pj.x, ey = pi.y - pj.y, ez = pi.z - pj.z;
            const T sq = sqlen(ex, ey, ez);
            const T r = sqrt(sq);
            if (r > c.h) return;
            const bool sym = pri && xj.proper;
            const T rbj = sym ? -xj.lb * c.inv_rho0 / xj.S : (T)0;
            const T sbj = sym ? -xj.lam * xj.lb / xj.S : (T)0;
            const T dw = (rb + rbj) * dwpoly(c, sq);
            T dx = dw * ex, dy = dw * ey, dz = dw * ez;
            if (r > (T)0) {
                const T s = gspiky(c, r);
                const T ss = 2 * (sb + sbj) * s * c.inv_rho0;
                const T gx = ss * ex + 2 * (sb * Gi.x - sbj * xj.G.x), gy = ss * ey + 2 * (sb * Gi.y - sbj * xj.G.y),
                        gz = ss * ez + 2 * (sb * Gi.z - sbj * xj.G.z);
                T jx, jy, jz;
                jspiky(c, r, ex, ey, ez, gx, gy, gz, jx, jy, jz);
                dx += c.inv_rho0 * jx;
                dy += c.inv_rho0 * jy;
                dz += c.inv_rho0 * jz;
            }
            pbx += dx;
            pby += dy;
            pbz += dz;
            if (!sym && j < c.n) atomic_add3(pbar, j, -dx, -dy, -dz);
        });
    pbx = group_sum(pbx);
    pby = group_sum(pby);
    pbz = group_sum(pbz);
    if (sub == 0) atomic_add3(pbar, i, pbx, pby, pbz);
}


// gather and predict reversed: xsbar[perm[g]] = pbar0[g];
// xbar_old += xsbar ; vbar_old = dt xsbar (v enters the step only through x*).
template <class T>
__global__ __launch_bounds__(AB) void kr_predict(AConsts<T> c, const uint32_t *__restrict__ perm,
                                                 const T4<T> *__restrict__ pbar0, T4<T> *__restrict__ xbar_old,
                                                 T4<T> *__restrict__ vbar_old) {
    const int32_t g = blockIdx.x * AB + threadIdx.x;
    if (g >= c.n) return;
    const uint32_t id = perm[g];
    const T4<T> b = pbar0[g];
    T4<T> x = xbar_old[id];
    x.x += b.x;
    x.y += b.y;
    x.z += b.z;
    xbar_old[id] = x;
    vbar_old[id] = T4<T>{c.dt * b.x, c.dt * b.y, c.dt * b.z, 0};
}

// State to dual numbers (derivative from dx, or zero) and back.
__global__ __launch_bounds__(AB) void ka_to_dual(int32_t n, const T4<double> *__restrict__ x,
                                                 const T4<double> *__restrict__ dx, T4<Dual> *__restrict__ out) {
    const int32_t i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    const T4<double> a = x[i];
    const T4<double> b = dx ? dx[i] : T4<double>{0.0, 0.0, 0.0, 0.0};
    out[i] = T4<Dual>{Dual(a.x, b.x), Dual(a.y, b.y), Dual(a.z, b.z), Dual(a.w, 0.0)};
}
__global__ __launch_bounds__(AB) void ka_from_dual(int32_t n, const T4<Dual> *__restrict__ in,
                                                   T4<double> *__restrict__ val, T4<double> *__restrict__ der) {
    const int32_t i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    const T4<Dual> a = in[i];
    val[i] = T4<double>{a.x.v, a.y.v, a.z.v, 0.0};
    der[i] = T4<double>{a.x.d, a.y.d, a.z.d, 0.0};
}

template <class T>
__global__ __launch_bounds__(AB) void ka_zero(T *__restrict__ p, size_t count) {
    for (size_t k = (size_t)blockIdx.x * AB + threadIdx.x; k < count; k += (size_t)gridDim.x * AB) p[k] = (T)0;
}

// The accumulators of one reverse step cleared in one launch (per particle).
template <class T>
__global__ __launch_bounds__(AB) void kr_clear(int32_t n, T4<T> *__restrict__ a, T4<T> *__restrict__ b,
                                               T4<T> *__restrict__ v, T *__restrict__ o, T *__restrict__ la,
                                               T *__restrict__ lb) {
    const int32_t i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    const T4<T> z{0, 0, 0, 0};
    a[i] = z;
    b[i] = z;
    v[i] = z;
    o[i] = (T)0;
    la[i] = (T)0;
    lb[i] = (T)0;
}

unsigned ablocks(size_t n) { return (unsigned)((n + AB - 1) / AB); }
unsigned zblocks(size_t n) {
    const size_t b = (n + AB - 1) / AB;
    return (unsigned)(b > 4096 ? 4096 : (b ? b : 1));
}

// ------------------------------------------------------------------ host
struct AdjBase {
    virtual ~AdjBase() {}
    virtual int forward(const void *pos4, const void *vel4, int host) = 0;
    virtual int positions(void *pos4, int host) = 0;
    virtual int velocities(void *vel4, int host) = 0;
    virtual int backward(const void *gpos4, int host, void *gvel4, double *gc, double *geps) = 0;
    virtual int replay(const void *pos4, const void *vel4, int host, void *pos_out4, uint64_t *clamp_flips) = 0;
    virtual int jvp(const void *pos4, const void *vel4, const void *dvel4, double dc, double de, int host,
                    void *pos_out4, void *dpos_out4) = 0;
    virtual void set_coefficients(double c, double eps) = 0;
    virtual void set_stream(hipStream_t st) = 0;
    virtual hipStream_t stream() const = 0;
};

template <class T>
struct StepTape {
    const uint32_t *perm = nullptr;  // the step's structure (kept by the forward pass)
    const int32_t *nbr9 = nullptr;
    T4<T> *p = nullptr;      // (I + 1) x n positions of the iterations (sorted)
    T *lam = nullptr, *sden = nullptr;  // I x n
    T4<T> *G = nullptr;      // I x n
    uint8_t *mask = nullptr; // I x n
    const uint8_t *proper = nullptr;  // n: the step's particles whose rows are exact (ka_proper)
    T4<T> *vs = nullptr, *vx = nullptr, *om = nullptr, *eta = nullptr;
    T *o = nullptr;
    // pair lists of the FP32 step ((I + 1) lists: the iterations', then vorticity's)
    uint16_t *ents = nullptr, *cnts = nullptr;
    const uint16_t *list_entries(int32_t k, uint32_t n) const { return ents ? ents + (size_t)k * kListRows * n : nullptr; }
    const uint16_t *list_counts(int32_t k, uint32_t n) const { return cnts ? cnts + (size_t)k * n : nullptr; }
};

template <class T>
struct Adj : AdjBase {
    uint32_t n;
    int32_t grid[3];
    size_t cells;
    int key_bits;
    int32_t steps, every, iters;
    pbf_params params;
    SolverConsts fc;         // float constants of the structure kernels
    AConsts<T> c;
    hipStream_t own = nullptr, st = nullptr;
    // structure machinery (float)
    float4 *xs_f = nullptr, *rec_f = nullptr;
    uint32_t *keys = nullptr;
    RadixScratch rs;
    int32_t *start = nullptr;
    unsigned long long *end64 = nullptr;
    uint32_t *epoch = nullptr;
    // state and checkpoints
    T4<T> *x = nullptr, *v = nullptr, *xf = nullptr, *vf = nullptr;  // working and final state
    std::vector<T4<T> *> ck_x, ck_v;
    std::vector<unsigned long long *> ck_end;
    std::vector<uint32_t *> ck_epoch;
    // structure of every step (kept on request, for replays)
    uint32_t *all_perm = nullptr;
    int32_t *all_nbr = nullptr;
    uint8_t *all_clamp = nullptr;            // clamp states of every step and iteration of the forward
    uint8_t *all_proper = nullptr;           // proper flags of every step (ka_proper)
    uint8_t *cur_proper = nullptr;           // where structure() writes them (null: not kept)
    uint32_t *skeys = nullptr;               // keys of the sorted records
    int32_t *tstart = nullptr, *tend = nullptr;  // true cell ranges
    unsigned long long *flips = nullptr;     // clamp decisions a frozen replay overrode
    bool have_all = false;
    // tape of one segment, and the segment it holds (-1: none)
    std::vector<StepTape<T>> tape;
    int32_t tape_segment = -1;
    T4<T> *xs = nullptr;
    T *lam_tmp = nullptr;
    uint2 *qpos = nullptr;  // quantized positions of the FP32 step's current records
    // reverse buffers
    T4<T> *xb = nullptr, *vb = nullptr, *xb2 = nullptr, *vb2 = nullptr, *pb_a = nullptr, *pb_b = nullptr;
    T4<T> *vxb = nullptr, *omb = nullptr, *vsb = nullptr, *etab = nullptr;
    T *ob = nullptr, *lb = nullptr, *lb2 = nullptr;
    double *acc = nullptr;
    bool forwarded = false;
    // pair lists and neighbour records of the plain step's kernels (FP32 handles)
    uint16_t *list_entries = nullptr, *list_count = nullptr;
    float4 *nrec = nullptr;
    // dual-number state of pbf_adjoint_jvp (FP64 handles, allocated on first use)
    T4<Dual> *jx = nullptr, *jv = nullptr, *jxs = nullptr, *jpa = nullptr, *jpb = nullptr, *jvs = nullptr,
             *jvx = nullptr, *jom = nullptr;
    Dual *jlam = nullptr, *jo = nullptr;

    ~Adj() override { release(); }

    void release() {
        dfree(xs_f);
        dfree(rec_f);
        dfree(keys);
        radix_free(rs);
        dfree(start);
        dfree(end64);
        dfree(epoch);
        dfree(x);
        dfree(v);
        dfree(xf);
        dfree(vf);
        for (auto &p : ck_x) dfree(p);
        for (auto &p : ck_v) dfree(p);
        for (auto &p : ck_end) dfree(p);
        for (auto &p : ck_epoch) dfree(p);
        dfree(all_perm);
        dfree(all_nbr);
        dfree(all_clamp);
        dfree(all_proper);
        dfree(skeys);
        dfree(tstart);
        dfree(tend);
        dfree(etab);
        dfree(flips);
        for (auto &t : tape) {
            dfree(t.p);
            dfree(t.lam);
            dfree(t.sden);
            dfree(t.G);
            dfree(t.mask);
            dfree(t.vs);
            dfree(t.vx);
            dfree(t.om);
            dfree(t.o);
            dfree(t.eta);
            dfree(t.ents);
            dfree(t.cnts);
        }
        dfree(xs);
        dfree(lam_tmp);
        dfree(qpos);
        dfree(xb);
        dfree(vb);
        dfree(xb2);
        dfree(vb2);
        dfree(pb_a);
        dfree(pb_b);
        dfree(vxb);
        dfree(omb);
        dfree(vsb);
        dfree(ob);
        dfree(lb);
        dfree(lb2);
        dfree(acc);
        dfree(list_entries);
        dfree(list_count);
        dfree(nrec);
        dfree(jx);
        dfree(jv);
        dfree(jxs);
        dfree(jpa);
        dfree(jpb);
        dfree(jvs);
        dfree(jvx);
        dfree(jom);
        dfree(jlam);
        dfree(jo);
        if (own) (void)hipStreamDestroy(own);
        own = nullptr;
    }

    hipError_t init(uint32_t n_, const int32_t g[3], const pbf_params &p, int32_t steps_, int32_t every_) {
        n = n_;
        for (int a = 0; a < 3; a++) grid[a] = g[a];
        cells = (size_t)g[0] * g[1] * g[2];
        key_bits = key_bits_of(grid);
        steps = steps_;
        every = every_;
        params = p;
        iters = (int32_t)p.iterations;
        fc = make_consts(n, grid, p, key_bits);
        set_consts(p.xsph_viscosity_c, p.vorticity_epsilon);
        hipError_t e;
        if ((e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking))) return e;
        st = own;
        if ((e = dalloc(&xs_f, n)) || (e = dalloc(&rec_f, n)) || (e = dalloc(&keys, n))) return e;
        if (!radix_alloc(rs, (int32_t)n)) return hipErrorOutOfMemory;
        if ((e = dalloc(&start, cells)) || (e = dalloc(&end64, cells)) || (e = dalloc(&epoch, 1))) return e;
        if ((e = dalloc(&x, n)) || (e = dalloc(&v, n)) || (e = dalloc(&xf, n)) || (e = dalloc(&vf, n))) return e;
        const int32_t nck = (steps + every - 1) / every;
        ck_x.assign(nck, nullptr);
        ck_v.assign(nck, nullptr);
        ck_end.assign(nck, nullptr);
        ck_epoch.assign(nck, nullptr);
        for (int32_t k = 0; k < nck; k++) {
            if ((e = dalloc(&ck_x[k], n)) || (e = dalloc(&ck_v[k], n)) || (e = dalloc(&ck_end[k], cells)) ||
                (e = dalloc(&ck_epoch[k], 1)))
                return e;
        }
        if ((e = dalloc(&all_perm, (size_t)steps * n)) || (e = dalloc(&all_nbr, (size_t)steps * 9 * n))) return e;
        if ((e = dalloc(&all_clamp, (size_t)steps * (iters > 0 ? iters : 1) * n)) || (e = dalloc(&flips, 1)))
            return e;
        if ((e = dalloc(&all_proper, (size_t)steps * n)) || (e = dalloc(&skeys, n)) || (e = dalloc(&tstart, cells)) ||
            (e = dalloc(&tend, cells)) || (e = dalloc(&etab, n)))
            return e;
        tape.assign(every, StepTape<T>());
        for (auto &t : tape) {
            if ((e = dalloc(&t.p, (size_t)(iters + 1) * n)) || (e = dalloc(&t.lam, (size_t)iters * n)) ||
                (e = dalloc(&t.sden, (size_t)iters * n)) || (e = dalloc(&t.G, (size_t)iters * n)) ||
                (e = dalloc(&t.mask, (size_t)iters * n)) || (e = dalloc(&t.vs, n)) || (e = dalloc(&t.vx, n)) ||
                (e = dalloc(&t.om, n)) || (e = dalloc(&t.o, n)) || (e = dalloc(&t.eta, n)))
                return e;
            if (std::is_same<T, float>::value &&
                ((e = dalloc(&t.ents, (size_t)(iters + 1) * kListRows * n)) ||
                 (e = dalloc(&t.cnts, (size_t)(iters + 1) * n))))
                return e;
        }
        if ((e = dalloc(&xs, n)) || (e = dalloc(&lam_tmp, n)) || (e = dalloc(&qpos, n))) return e;
        if ((e = dalloc(&xb, n)) || (e = dalloc(&vb, n)) || (e = dalloc(&xb2, n)) || (e = dalloc(&vb2, n)) ||
            (e = dalloc(&pb_a, n)) || (e = dalloc(&pb_b, n)) || (e = dalloc(&vxb, n)) || (e = dalloc(&omb, n)) ||
            (e = dalloc(&vsb, n)) || (e = dalloc(&ob, n)) || (e = dalloc(&lb, n)) || (e = dalloc(&lb2, n)) ||
            (e = dalloc(&acc, 2)))
            return e;
        if (std::is_same<T, float>::value &&
            ((e = dalloc(&list_entries, (size_t)kListRows * n)) || (e = dalloc(&list_count, n)) ||
             (e = dalloc(&nrec, n))))
            return e;
        return hipSuccess;
    }

    void set_consts(double xsph, double veps) {
        const pbf_params &p = params;
        c.n = (int32_t)n;
        c.gx = grid[0];
        c.gy = grid[1];
        c.gz = grid[2];
        const T h = (T)p.h;
        c.h = h;
        c.hh = h * h;
        c.h6 = h * h * h * h * h * h;
        c.h9 = h * h * h * h * h * h * h * h * h;
        c.inv_rho0 = (T)p.one_over_rho_0;
        c.eps = (T)p.epsilon;
        c.gravity = (T)p.gravity;
        c.dt = (T)p.timestep;
        c.k = (T)p.tensile_instability_k;
        c.scale = (T)p.tensile_instability_scale;
        c.xsph_c = (T)xsph;
        c.vort_eps = (T)veps;
        for (int a = 0; a < 3; a++) {
            c.lo[a] = (T)p.wall_lo[a];
            c.hi[a] = (T)p.wall_hi[a];
            c.gdir[a] = (T)p.gravity_dir[a];
        }
        c.extforce = p.extforce;
        c.vorticity = p.vorticity;
        fc.xsph_c = (float)xsph;  // the plain kernels' constants (FP32 forward)
        fc.vort_eps = (float)veps;
    }
    // the tape was made with the previous coefficients: the next backward pass
    // recomputes every segment
    void set_coefficients(double cc, double eps) override {
        set_consts(cc, eps);
        tape_segment = -1;
    }
    // work queued on the old stream finishes first, so calls stay in order
    void set_stream(hipStream_t s) override {
        const hipStream_t next = s ? s : own;
        if (next != st) (void)hipStreamSynchronize(st);
        st = next;
    }
    hipStream_t stream() const override { return st; }

    template <class U>
    void zero(U *p, size_t count) {
        ka_zero<U><<<zblocks(count), AB, 0, st>>>(p, count);
    }

    // Structure of the step from state (x, v): predicted records, sort order
    // and neighbour table, by the plain step's kernels (float), into perm/nbr9.
    void structure(uint32_t *perm, int32_t *nbr9) {
        launch_keys_of_records(fc, xs_f, keys, st);
        radix_sort_perm(rs, keys, (int32_t)n, key_bits, perm, st);
        launch_gather_records((int32_t)n, perm, xs_f, rec_f, st);
        launch_begin_cells(start, cells, epoch, st);
        launch_findcells(fc, rec_f, start, end64, epoch, st);
        launch_neighbourcells(fc, rec_f, start, end64, nbr9, st);
        if (cur_proper) {
            launch_keys_of_records(fc, rec_f, skeys, st);
            launch_fill_i32(tstart, -1, cells, st);
            const unsigned nb = ablocks(n);
            ka_true_cells<<<nb, AB, 0, st>>>((int32_t)n, grid[0], grid[1], grid[2], skeys, tstart, tend);
            ka_proper<<<nb, AB, 0, st>>>((int32_t)n, grid[0], grid[1], grid[2], cells < ((size_t)1 << 24) ? 1 : 0,
                                         rec_f, skeys, tstart, tend, nbr9, cur_proper);
        }
    }

    // One step on (x, v) in place. With a tape, everything the reverse pass
    // reads is kept; perm/nbr9 given: fixed structure, else computed into them.
    // clamp: the step's clamp states (iters x n) to record, or with frozen,
    // the recorded ones to impose (counting overridden decisions in flips).
    void step(StepTape<T> *tp, uint32_t *perm, int32_t *nbr9, bool compute_structure, uint8_t *clamp = nullptr,
              bool frozen = false) {
        if constexpr (std::is_same<T, float>::value) {
            if (!frozen) return step_plain(tp, perm, nbr9, compute_structure, clamp);
        }
        const unsigned nb = ablocks(n);
        ka_predict<T><<<nb, AB, 0, st>>>(c, x, v, xs, xs_f);
        if (compute_structure) structure(perm, nbr9);
        T4<T> *p0 = tp ? tp->p : pb_a;  // pb_a/pb_b double as iteration buffers without a tape
        ka_gather<T><<<nb, AB, 0, st>>>((int32_t)n, perm, xs, p0);
        T4<T> *cur = p0;
        for (int32_t it = 0; it < iters; it++) {
            T4<T> *nxt = tp ? tp->p + (size_t)(it + 1) * n : (cur == pb_a ? pb_b : pb_a);
            T *lam = tp ? tp->lam + (size_t)it * n : lam_tmp;
            ka_lambda<T><<<nb, AB, 0, st>>>(c, cur, nbr9, lam, tp ? tp->sden + (size_t)it * n : nullptr,
                                            tp ? tp->G + (size_t)it * n : nullptr);
            uint8_t *cl = clamp ? clamp + (size_t)it * n : nullptr;
            ka_updatepos<T><<<nb, AB, 0, st>>>(c, cur, lam, nbr9, nxt,
                                               tp ? tp->mask + (size_t)it * n : (frozen ? nullptr : cl),
                                               frozen ? cl : nullptr, flips);
            cur = nxt;
        }
        T4<T> *vs = tp ? tp->vs : vsb;
        ka_update<T><<<nb, AB, 0, st>>>(c, perm, cur, x, v, vs);
        if (c.vorticity) {
            T4<T> *vx = tp ? tp->vx : vxb;
            T4<T> *om = tp ? tp->om : omb;
            T *o = tp ? tp->o : ob;
            // a zero record's velocity is that of id 0 (imageLoad (velocitytexture, 0))
            ka_vort_a<T><<<nb, AB, 0, st>>>(c, cur, nbr9, vs, v, vx, om, o);
            ka_vort_b<T><<<nb, AB, 0, st>>>(c, perm, cur, nbr9, vx, om, o, v, tp ? tp->eta : nullptr);
        }
    }

    // The FP32 step on the plain step's kernels (pbf_kernels.hip: candidate
    // scans with pair lists, the closing update in the last position update),
    // whose results equal the forward kernels' bit for bit; the tape values
    // come out of them as extra outputs.
    void step_plain(StepTape<T> *tp, uint32_t *perm, int32_t *nbr9, bool compute_structure, uint8_t *clamp) {
        const unsigned nb = ablocks(n);
        ka_predict<T><<<nb, AB, 0, st>>>(c, x, v, xs, xs_f);
        if (compute_structure) structure(perm, nbr9);
        float4 *p0 = reinterpret_cast<float4 *>(tp ? tp->p : pb_a);
        launch_gather_records((int32_t)n, perm, xs_f, p0, st);
        launch_quantize(fc, p0, qpos, (int32_t)n, st);
        PairList pl;
        pl.entries = list_entries;
        pl.count = list_count;
        pl.cap = kListCap;
        float4 *vs = reinterpret_cast<float4 *>(tp ? tp->vs : vsb);
        StepUpdate up;
        up.pos = reinterpret_cast<float4 *>(x);
        up.vel = reinterpret_cast<float4 *>(v);
        up.vel_sorted = vs;
        float4 *cur = p0;
        for (int32_t it = 0; it < iters; it++) {
            float4 *nxt = reinterpret_cast<float4 *>(tp ? tp->p + (size_t)(it + 1) * n
                                                        : (cur == reinterpret_cast<float4 *>(pb_a) ? pb_b : pb_a));
            float *lam = reinterpret_cast<float *>(tp ? tp->lam + (size_t)it * n : lam_tmp);
            SolverTape tl, tu;
            if (tp) {
                tl.sden = reinterpret_cast<float *>(tp->sden + (size_t)it * n);
                tl.G = reinterpret_cast<float4 *>(tp->G + (size_t)it * n);
                tu.clamp = tp->mask + (size_t)it * n;
            } else if (clamp) {
                tu.clamp = clamp + (size_t)it * n;
            }
            if (tp) {  // the reverse pass reads each iteration's list
                pl.entries = tp->ents + (size_t)it * kListRows * n;
                pl.count = tp->cnts + (size_t)it * n;
            }
            launch_lambda(fc, cur, qpos, nbr9, lam, nrec, pl, st, false, nullptr, &tl);
            launch_updatepos(fc, cur, lam, nrec, nbr9, pl, nxt, qpos, st, false, nullptr,
                             it + 1 == iters ? &up : nullptr, &tu);
            cur = nxt;
        }
        if (iters == 0) launch_update(fc, cur, up.pos, up.vel, vs, st);
        if (c.vorticity) {
            float4 *vx = reinterpret_cast<float4 *>(tp ? tp->vx : vxb);
            float4 *om = reinterpret_cast<float4 *>(tp ? tp->om : omb);
            float *o = reinterpret_cast<float *>(tp ? tp->o : ob);
            SolverTape te;
            if (tp) {
                te.eta = reinterpret_cast<float4 *>(tp->eta);
                pl.entries = tp->ents + (size_t)iters * kListRows * n;
                pl.count = tp->cnts + (size_t)iters * n;
            }
            launch_vorticity_a(fc, cur, qpos, nbr9, vs, up.vel, vx, om, o, nrec, pl, st, false);
            launch_vorticity_b(fc, cur, nbr9, pl, vx, om, nrec, up.vel, st, false, nullptr, &te);
        }
    }

    hipError_t save_ck(int32_t k) {
        hipError_t e;
        if ((e = hipMemcpyAsync(ck_x[k], x, n * sizeof(T4<T>), hipMemcpyDeviceToDevice, st))) return e;
        if ((e = hipMemcpyAsync(ck_v[k], v, n * sizeof(T4<T>), hipMemcpyDeviceToDevice, st))) return e;
        if ((e = hipMemcpyAsync(ck_end[k], end64, cells * 8, hipMemcpyDeviceToDevice, st))) return e;
        return hipMemcpyAsync(ck_epoch[k], epoch, 4, hipMemcpyDeviceToDevice, st);
    }
    hipError_t load_ck(int32_t k) {
        hipError_t e;
        if ((e = hipMemcpyAsync(x, ck_x[k], n * sizeof(T4<T>), hipMemcpyDeviceToDevice, st))) return e;
        if ((e = hipMemcpyAsync(v, ck_v[k], n * sizeof(T4<T>), hipMemcpyDeviceToDevice, st))) return e;
        if ((e = hipMemcpyAsync(end64, ck_end[k], cells * 8, hipMemcpyDeviceToDevice, st))) return e;
        return hipMemcpyAsync(epoch, ck_epoch[k], 4, hipMemcpyDeviceToDevice, st);
    }

    hipError_t upload(T4<T> *dx, T4<T> *dv, const void *pos4, const void *vel4, int host) {
        const hipMemcpyKind kd = host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
        hipError_t e;
        if ((e = hipMemcpyAsync(dx, pos4, n * sizeof(T4<T>), kd, st))) return e;
        if ((e = hipMemcpyAsync(dv, vel4, n * sizeof(T4<T>), kd, st))) return e;
        return host ? hipStreamSynchronize(st) : hipSuccess;
    }

    int forward(const void *pos4, const void *vel4, int host) override {
        PBF_HIP(upload(x, v, pos4, vel4, host));
        // a fresh end table, as a new SPH handle's
        PBF_HIP(hipMemsetAsync(end64, 0, cells * 8, st));
        PBF_HIP(hipMemsetAsync(epoch, 0, 4, st));
        // every step's structure is kept: the reverse pass recomputes the
        // states and tapes of a segment from its checkpoint with it; the last
        // segment, which the reverse pass starts with, is taped here
        const int32_t nck = (steps + every - 1) / every, last0 = (nck - 1) * every;
        tape_segment = -1;
        for (int32_t s = 0; s < steps; s++) {
            if (s % every == 0) PBF_HIP(save_ck(s / every));
            uint32_t *perm = all_perm + (size_t)s * n;
            int32_t *nbr = all_nbr + (size_t)s * 9 * n;
            uint8_t *clamp = all_clamp + (size_t)s * iters * n;
            cur_proper = all_proper + (size_t)s * n;
            if (s < last0) {
                step(nullptr, perm, nbr, true, clamp);
                continue;
            }
            StepTape<T> &t = tape[s - last0];
            t.perm = perm;
            t.nbr9 = nbr;
            t.proper = cur_proper;
            step(&t, perm, nbr, true);
            // the tape's clamp states are the step's (frozen replays read them)
            if (iters > 0) PBF_HIP(hipMemcpyAsync(clamp, t.mask, (size_t)iters * n, hipMemcpyDeviceToDevice, st));
        }
        cur_proper = nullptr;
        tape_segment = nck - 1;
        PBF_HIP(hipMemcpyAsync(xf, x, n * sizeof(T4<T>), hipMemcpyDeviceToDevice, st));
        PBF_HIP(hipMemcpyAsync(vf, v, n * sizeof(T4<T>), hipMemcpyDeviceToDevice, st));
        have_all = true;
        forwarded = true;
        PBF_HIP(hipGetLastError());
        return PBF_OK;
    }

    int positions(void *pos4, int host) override {
        PBF_HIP(hipMemcpyAsync(pos4, xf, n * sizeof(T4<T>), host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                               st));
        if (host) PBF_HIP(hipStreamSynchronize(st));
        return PBF_OK;
    }
    int velocities(void *vel4, int host) override {
        PBF_HIP(hipMemcpyAsync(vel4, vf, n * sizeof(T4<T>), host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                               st));
        if (host) PBF_HIP(hipStreamSynchronize(st));
        return PBF_OK;
    }

    // Reverse pass of one step from its tape: (xb, vb) after the step -> (xb, vb) before it.
    void reverse_step(const StepTape<T> &t) {
        const unsigned nb = ablocks(n);
        const unsigned ng = ablocks((size_t)kAdjLanes * n);  // a lane group per particle
        const T4<T> *pI = t.p + (size_t)iters * n;
        // every accumulator the first iteration needs; each iteration's lambda
        // launch clears the next one's
        kr_clear<T><<<nb, AB, 0, st>>>((int32_t)n, pb_a, pb_b, vsb, ob, lb, lb2);
        const T4<T> *vsbar = nullptr;
        if (c.vorticity) {
            kr_vort_b_own<T><<<nb, AB, 0, st>>>(c, t.perm, t.om, t.eta, vb, vxb, omb, etab, acc);
            kr_vort_b<T><<<ng, AB, 0, st>>>(c, pI, t.nbr9, t.list_entries(iters, n), t.list_counts(iters, n),
                                            t.proper, t.o, etab, ob, pb_a);
            kr_omega_norm<T><<<nb, AB, 0, st>>>((int32_t)n, t.om, t.o, ob, omb);
            kr_vort_a<T><<<ng, AB, 0, st>>>(c, pI, t.nbr9, t.list_entries(iters, n), t.list_counts(iters, n),
                                            t.proper, t.vs, vxb, omb, vsb, pb_a, acc);
            vsbar = vsb;
        }
        kr_update<T><<<nb, AB, 0, st>>>(c, t.perm, xb, vb, vsbar, pb_a, xb2);
        T4<T> *next = pb_a, *prev = pb_b;
        T *lcur = lb, *lnext = lb2;
        for (int32_t it = iters - 1; it >= 0; it--) {
            const T4<T> *p = t.p + (size_t)it * n;
            const T *lam = t.lam + (size_t)it * n;
            kr_updatepos<T><<<ng, AB, 0, st>>>(c, p, lam, t.nbr9, t.list_entries(it, n), t.list_counts(it, n),
                                               t.proper, t.mask + (size_t)it * n, next, prev, lcur);
            kr_lambda<T><<<ng, AB, 0, st>>>(c, p, t.nbr9, t.list_entries(it, n), t.list_counts(it, n), t.proper, lam,
                                            t.sden + (size_t)it * n, t.G + (size_t)it * n, lcur, prev, next, lnext);
            std::swap(next, prev);
            std::swap(lcur, lnext);
        }
        kr_predict<T><<<nb, AB, 0, st>>>(c, t.perm, next, xb2, vb2);
        std::swap(xb, xb2);
        std::swap(vb, vb2);
    }

    int backward(const void *gpos4, int host, void *gvel4, double *gc, double *geps) override {
        if (!forwarded) return invalid("pbf_adjoint_backward: no forward rollout");
        PBF_HIP(hipMemcpyAsync(xb, gpos4, n * sizeof(T4<T>), host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                               st));
        zero(reinterpret_cast<T *>(vb), 4 * (size_t)n);
        zero(acc, 2);
        const int32_t nck = (steps + every - 1) / every;
        for (int32_t k = nck - 1; k >= 0; k--) {
            const int32_t s0 = k * every, s1 = std::min(steps, s0 + every);
            if (tape_segment != k) {
                PBF_HIP(load_ck(k));
                for (int32_t s = s0; s < s1; s++) {
                    StepTape<T> &t = tape[s - s0];
                    uint32_t *perm = all_perm + (size_t)s * n;
                    int32_t *nbr = all_nbr + (size_t)s * 9 * n;
                    t.perm = perm;
                    t.nbr9 = nbr;
                    t.proper = all_proper + (size_t)s * n;
                    step(&t, perm, nbr, false);
                }
                tape_segment = k;
            }
            for (int32_t s = s1 - 1; s >= s0; s--) reverse_step(tape[s - s0]);
        }
        PBF_HIP(hipGetLastError());
        if (gvel4)
            PBF_HIP(hipMemcpyAsync(gvel4, vb, n * sizeof(T4<T>), host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                                   st));
        double h_acc[2] = {0.0, 0.0};
        PBF_HIP(hipMemcpyAsync(h_acc, acc, sizeof h_acc, hipMemcpyDeviceToHost, st));
        PBF_HIP(hipStreamSynchronize(st));
        if (gc) *gc = h_acc[0];
        if (geps) *geps = h_acc[1];
        return PBF_OK;
    }

    int replay(const void *pos4, const void *vel4, int host, void *pos_out4, uint64_t *clamp_flips) override {
        if (!have_all) return invalid("pbf_adjoint_replay: no forward rollout");
        PBF_HIP(upload(x, v, pos4, vel4, host));
        const bool frozen = clamp_flips != nullptr;
        if (frozen) zero(flips, 1);
        for (int32_t s = 0; s < steps; s++)
            step(nullptr, all_perm + (size_t)s * n, all_nbr + (size_t)s * 9 * n, false,
                 frozen ? all_clamp + (size_t)s * iters * n : nullptr, frozen);
        PBF_HIP(hipGetLastError());
        PBF_HIP(hipMemcpyAsync(pos_out4, x, n * sizeof(T4<T>), host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                               st));
        if (frozen) {
            unsigned long long h = 0;
            PBF_HIP(hipMemcpyAsync(&h, flips, sizeof h, hipMemcpyDeviceToHost, st));
            PBF_HIP(hipStreamSynchronize(st));
            *clamp_flips = h;
        }
        if (host) PBF_HIP(hipStreamSynchronize(st));
        return PBF_OK;
    }

    // Directional derivative of the replay map (the structure of the last
    // forward, the reverse pass's clamp and eta = 0 conventions) along
    // (dvel4, dc, de), by the forward kernels on dual numbers.
    int jvp(const void *pos4, const void *vel4, const void *dvel4, double dc, double de, int host, void *pos_out4,
            void *dpos_out4) override {
        if constexpr (!std::is_same<T, double>::value) {
            return invalid("pbf_adjoint_jvp: FP64 handles only");
        } else {
            if (!have_all) return invalid("pbf_adjoint_jvp: no forward rollout");
            hipError_t e = hipSuccess;
            if (!jx && ((e = dalloc(&jx, n)) || (e = dalloc(&jv, n)) || (e = dalloc(&jxs, n)) ||
                        (e = dalloc(&jpa, n)) || (e = dalloc(&jpb, n)) || (e = dalloc(&jvs, n)) ||
                        (e = dalloc(&jvx, n)) || (e = dalloc(&jom, n)) || (e = dalloc(&jlam, n)) ||
                        (e = dalloc(&jo, n))))
                return hip_fail(e, "pbf_adjoint_jvp");
            const hipMemcpyKind kd = host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
            PBF_HIP(hipMemcpyAsync(x, pos4, n * sizeof(T4<T>), kd, st));
            PBF_HIP(hipMemcpyAsync(v, vel4, n * sizeof(T4<T>), kd, st));
            PBF_HIP(hipMemcpyAsync(xb, dvel4, n * sizeof(T4<T>), kd, st));
            const unsigned nb = ablocks(n);
            ka_to_dual<<<nb, AB, 0, st>>>((int32_t)n, x, nullptr, jx);
            ka_to_dual<<<nb, AB, 0, st>>>((int32_t)n, v, xb, jv);
            AConsts<Dual> d;
            d.n = c.n;
            d.gx = c.gx;
            d.gy = c.gy;
            d.gz = c.gz;
            d.h = c.h;
            d.hh = c.hh;
            d.h6 = c.h6;
            d.h9 = c.h9;
            d.inv_rho0 = c.inv_rho0;
            d.eps = c.eps;
            d.gravity = c.gravity;
            d.dt = c.dt;
            d.k = c.k;
            d.scale = c.scale;
            d.xsph_c = Dual(c.xsph_c, dc);
            d.vort_eps = Dual(c.vort_eps, de);
            for (int a = 0; a < 3; a++) {
                d.lo[a] = c.lo[a];
                d.hi[a] = c.hi[a];
                d.gdir[a] = c.gdir[a];
            }
            d.extforce = c.extforce;
            d.vorticity = c.vorticity;
            for (int32_t s = 0; s < steps; s++) {
                const uint32_t *perm = all_perm + (size_t)s * n;
                const int32_t *nbr9 = all_nbr + (size_t)s * 9 * n;
                ka_predict<Dual><<<nb, AB, 0, st>>>(d, jx, jv, jxs, xs_f);
                ka_gather<Dual><<<nb, AB, 0, st>>>((int32_t)n, perm, jxs, jpa);
                T4<Dual> *cur = jpa, *nxt = jpb;
                for (int32_t it = 0; it < iters; it++) {
                    ka_lambda<Dual><<<nb, AB, 0, st>>>(d, cur, nbr9, jlam, nullptr, nullptr);
                    ka_updatepos<Dual><<<nb, AB, 0, st>>>(d, cur, jlam, nbr9, nxt, nullptr, nullptr, nullptr);
                    std::swap(cur, nxt);
                }
                ka_update<Dual><<<nb, AB, 0, st>>>(d, perm, cur, jx, jv, jvs);
                if (d.vorticity) {
                    ka_vort_a<Dual><<<nb, AB, 0, st>>>(d, cur, nbr9, jvs, jv, jvx, jom, jo);
                    ka_vort_b<Dual><<<nb, AB, 0, st>>>(d, perm, cur, nbr9, jvx, jom, jo, jv, nullptr);
                }
            }
            ka_from_dual<<<nb, AB, 0, st>>>((int32_t)n, jx, x, xb);
            PBF_HIP(hipGetLastError());
            const hipMemcpyKind ko = host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
            if (pos_out4) PBF_HIP(hipMemcpyAsync(pos_out4, x, n * sizeof(T4<T>), ko, st));
            PBF_HIP(hipMemcpyAsync(dpos_out4, xb, n * sizeof(T4<T>), ko, st));
            if (host) PBF_HIP(hipStreamSynchronize(st));
            return PBF_OK;
        }
    }
};

}  // namespace
}  // namespace pbf

using namespace pbf;

struct pbf_adjoint {
    AdjBase *impl = nullptr;
    int fp64 = 0;
};

extern "C" {

int pbf_adjoint_create(pbf_adjoint **out, uint32_t n, int32_t gx, int32_t gy, int32_t gz, const pbf_params *p,
                       int32_t steps, int32_t checkpoint_every, int32_t fp64) {
    if (!out) return invalid("null output handle");
    *out = nullptr;
    if (n == 0 || n % 256 != 0 || n >= (1u << 24)) return invalid("n must be a positive multiple of 256 below 2^24");
    if (!valid_grid(gx, gy, gz)) return invalid("bad grid");
    int rc = check_params(p);
    if (rc) return rc;
    if (steps < 1 || checkpoint_every < 1) return invalid("steps and checkpoint_every must be positive");
    if (pbf_device_count() < 1) {
        set_error("no HIP device");
        return PBF_ERR_NODEVICE;
    }
    const int32_t g[3] = {gx, gy, gz};
    pbf_adjoint *a = new (std::nothrow) pbf_adjoint();
    if (!a) return invalid("host allocation failed");
    a->fp64 = fp64 != 0;
    hipError_t e;
    if (a->fp64) {
        auto *impl = new (std::nothrow) Adj<double>();
        a->impl = impl;
        e = impl ? impl->init(n, g, *p, steps, std::min(checkpoint_every, steps)) : hipErrorOutOfMemory;
    } else {
        auto *impl = new (std::nothrow) Adj<float>();
        a->impl = impl;
        e = impl ? impl->init(n, g, *p, steps, std::min(checkpoint_every, steps)) : hipErrorOutOfMemory;
    }
    if (e != hipSuccess) {
        int r = hip_fail(e, "pbf_adjoint_create");
        delete a->impl;
        delete a;
        return r;
    }
    *out = a;
    return PBF_OK;
}

int pbf_adjoint_destroy(pbf_adjoint *a) {
    if (!a) return PBF_OK;
    delete a->impl;
    delete a;
    return PBF_OK;
}

int pbf_adjoint_set_coefficients(pbf_adjoint *a, double xsph_c, double vorticity_epsilon) {
    if (!a) return invalid("null adjoint handle");
    a->impl->set_coefficients(xsph_c, vorticity_epsilon);
    return PBF_OK;
}

int pbf_adjoint_forward(pbf_adjoint *a, const void *pos4, const void *vel4, int host) {
    if (!a || !pos4 || !vel4) return invalid("null argument");
    return a->impl->forward(pos4, vel4, host);
}

int pbf_adjoint_positions(pbf_adjoint *a, void *pos4, int host) {
    if (!a || !pos4) return invalid("null argument");
    return a->impl->positions(pos4, host);
}

int pbf_adjoint_velocities(pbf_adjoint *a, void *vel4, int host) {
    if (!a || !vel4) return invalid("null argument");
    return a->impl->velocities(vel4, host);
}

int pbf_adjoint_backward(pbf_adjoint *a, const void *grad_pos4, int host, void *grad_vel4, double *grad_xsph_c,
                         double *grad_vorticity_epsilon) {
    if (!a || !grad_pos4) return invalid("null argument");
    return a->impl->backward(grad_pos4, host, grad_vel4, grad_xsph_c, grad_vorticity_epsilon);
}

int pbf_adjoint_replay(pbf_adjoint *a, const void *pos4, const void *vel4, int host, void *pos_out4) {
    if (!a || !pos4 || !vel4 || !pos_out4) return invalid("null argument");
    return a->impl->replay(pos4, vel4, host, pos_out4, nullptr);
}

int pbf_adjoint_jvp(pbf_adjoint *a, const void *pos4, const void *vel4, const void *dvel4, double dxsph_c,
                    double dvorticity_epsilon, int host, void *pos_out4, void *dpos_out4) {
    if (!a || !pos4 || !vel4 || !dvel4 || !dpos_out4) return invalid("null argument");
    return a->impl->jvp(pos4, vel4, dvel4, dxsph_c, dvorticity_epsilon, host, pos_out4, dpos_out4);
}

int pbf_adjoint_replay_frozen(pbf_adjoint *a, const void *pos4, const void *vel4, int host, void *pos_out4,
                              uint64_t *clamp_flips) {
    if (!a || !pos4 || !vel4 || !pos_out4 || !clamp_flips) return invalid("null argument");
    return a->impl->replay(pos4, vel4, host, pos_out4, clamp_flips);
}

int pbf_adjoint_set_stream(pbf_adjoint *a, void *stream) {
    if (!a) return invalid("null adjoint handle");
    a->impl->set_stream((hipStream_t)stream);
    return PBF_OK;
}

void *pbf_adjoint_get_stream(const pbf_adjoint *a) { return a ? (void *)a->impl->stream() : nullptr; }

}  // extern "C"
