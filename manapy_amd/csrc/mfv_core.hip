// libmfv core: error plumbing, mesh checks, generic operators (ghost fill, face->cell,
// least-squares gradient, Barth-Jespersen limiter) and linear advection.
//
// Every kernel runs one thread per cell (or per listed face) and gathers what it needs;
// nothing is scattered and no atomics touch floating-point sums, so results are
// deterministic and independent of the launch shape.
#include <stdarg.h>

#include <mutex>
#include <vector>

#include "mfv_common.hpp"

namespace mfv {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return MFV_EHIP;
    }
    return MFV_OK;
}

int validate_mesh(const mfv_mesh *m, int need_lsq, int need_order_split, int need_order_sorted)
{
    MFV_REQUIRE(m != nullptr, "mesh descriptor is NULL");
    MFV_REQUIRE(m->dim == 2 || m->dim == 3, "mesh dim must be 2 or 3, got %d", m->dim);
    MFV_REQUIRE(m->nb_cells >= 0 && m->nb_faces >= 0, "negative mesh sizes");
    MFV_REQUIRE(m->max_cell_faces >= 1 && m->max_cell_faces <= 8, "max_cell_faces %d out of [1, 8]",
                m->max_cell_faces);
    MFV_REQUIRE(m->face_cells && m->face_name && m->face_normal && m->face_mesure && m->face_center &&
                    m->cell_center && m->cell_volume && m->cell_faces && m->cell_nfaces,
                "mesh descriptor has NULL geometry/connectivity pointers");
    if (need_lsq)
        MFV_REQUIRE(m->lsq_cells && m->lsq_ncells && m->lsq_ghosts && m->lsq_nghosts && m->face_ghost &&
                        m->lsq_width >= 0 && m->ghost_width >= 0,
                    "mesh descriptor lacks the least-squares stencil");
    if (need_order_split) MFV_REQUIRE(m->order_split, "mesh descriptor lacks order_split");
    if (need_order_sorted) MFV_REQUIRE(m->order_sorted, "mesh descriptor lacks order_sorted");
    return MFV_OK;
}

}  // namespace mfv

using namespace mfv;

extern "C" const char *mfv_version(void) { return "mfv 0.1 gfx950"; }
extern "C" const char *mfv_last_error(void) { return mfv::g_err; }

extern "C" int mfv_device_count(int *count_host)
{
    MFV_REQUIRE(count_host, "count_host is NULL");
    hipError_t e = hipGetDeviceCount(count_host);
    if (e != hipSuccess) {
        *count_host = 0;
        set_error("hipGetDeviceCount: %s", hipGetErrorString(e));
        return MFV_EHIP;
    }
    return MFV_OK;
}

// ---------------------------------------------------------------------------------------------
// ghost values (boundary/bc_utils.py:45-68), face->cell (ast/ast_utils.py:8-18)

__global__ void __launch_bounds__(kBlock) k_ghost_value(const int32_t *__restrict__ face_cells, int kind,
                                                         const double *__restrict__ src,
                                                         double *__restrict__ w_ghost,
                                                         const int32_t *__restrict__ faces, int n)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int f = faces[k];
    if (kind == 0) w_ghost[f] = src[face_cells[2 * f]];
    else if (kind == 1) w_ghost[f] = -1 * src[face_cells[2 * f]];
    else w_ghost[f] = src[f];
}

extern "C" int mfv_ghost_value(const mfv_mesh *m, int kind, const double *src, double *w_ghost,
                               const int32_t *faces, int nfaces, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(kind >= 0 && kind <= 2, "ghost kind %d not in {0,1,2}", kind);
    MFV_REQUIRE(nfaces >= 0, "nfaces < 0");
    if (nfaces == 0) return MFV_OK;
    MFV_REQUIRE(src && w_ghost && faces, "NULL array");
    hipLaunchKernelGGL(k_ghost_value, dim3(grid_for(nfaces)), dim3(kBlock), 0, as_stream(stream), m->face_cells,
                       kind, src, w_ghost, faces, nfaces);
    return check_launch("k_ghost_value");
}

__global__ void __launch_bounds__(kBlock) k_facetocell(Mesh m, const double *__restrict__ u_face,
                                                        double *__restrict__ u_c)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    int nf = m.cell_nfaces[i];
    double s = 0.;
    for (int j = 0; j < nf; j++) s += u_face[m.cell_faces[(size_t)j * m.C + i]];
    u_c[i] = s / (double)nf;
}

extern "C" int mfv_facetocell(const mfv_mesh *m, const double *u_face, double *u_c, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(u_face && u_c, "NULL array");
    hipLaunchKernelGGL(k_facetocell, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       to_kernel(m), u_face, u_c);
    return check_launch("k_facetocell");
}

// ---------------------------------------------------------------------------------------------
// least-squares cell gradient (ast/functions2d.py:5-104, functions3d.py:4-120)

template <int DIM>
__device__ __forceinline__ void lsq_solve(const double (&I)[6], const double (&J)[3], double &gx, double &gy,
                                          double &gz)
{
    // I = {xx, yy, xy, zz, xz, yz}
    if (DIM == 2) {
        double dia = I[0] * I[1] - I[2] * I[2];
        gx = (I[1] * J[0] - I[2] * J[1]) / dia;
        gy = (I[0] * J[1] - I[2] * J[0]) / dia;
        gz = 0.;
    } else {
        double xx = I[0], yy = I[1], xy = I[2], zz = I[3], xz = I[4], yz = I[5];
        double dia = xx * yy * zz + 2. * xy * xz * yz - xx * (yz * yz) - yy * (xz * xz) - zz * (xy * xy);
        gx = ((yy * zz - yz * yz) * J[0] + (xz * yz - xy * zz) * J[1] + (xy * yz - xz * yy) * J[2]) / dia;
        gy = ((xz * yz - xy * zz) * J[0] + (xx * zz - xz * xz) * J[1] + (xy * xz - yz * xx) * J[2]) / dia;
        gz = ((xy * yz - xz * yy) * J[0] + (xy * xz - yz * xx) * J[1] + (xx * yy - xy * xy) * J[2]) / dia;
    }
}

template <int DIM>
__device__ __forceinline__ void lsq_accumulate(double (&I)[6], double (&J)[3], double jx, double jy, double jz,
                                               double dw)
{
    I[0] += jx * jx;
    I[1] += jy * jy;
    if (DIM == 3) I[3] += jz * jz;
    I[2] += (jx * jy);
    if (DIM == 3) {
        I[4] += (jx * jz);
        I[5] += (jy * jz);
    }
    J[0] += (jx * dw);
    J[1] += (jy * dw);
    if (DIM == 3) J[2] += (jz * dw);
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_cell_gradient(Mesh m, const double *__restrict__ w_c,
                                                           const double *__restrict__ w_ghost,
                                                           double *__restrict__ wx, double *__restrict__ wy,
                                                           double *__restrict__ wz)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    const double cx = m.cell_center[3 * i], cy = m.cell_center[3 * i + 1], cz = m.cell_center[3 * i + 2];
    const double wi = w_c[i];
    double I[6] = {0., 0., 0., 0., 0., 0.}, J[3] = {0., 0., 0.};
    const int nn = m.lsq_ncells[i];
    for (int j = 0; j < nn; j++) {
        int c = m.lsq_cells[(size_t)j * m.C + i];
        const double *p = m.cell_center + 3 * (size_t)c;
        lsq_accumulate<DIM>(I, J, p[0] - cx, p[1] - cy, DIM == 3 ? p[2] - cz : 0., w_c[c] - wi);
    }
    const int ng = m.lsq_nghosts[i];
    for (int j = 0; j < ng; j++) {
        int f = m.lsq_ghosts[(size_t)j * m.C + i];
        const double *p = m.face_ghost + 3 * (size_t)f;
        lsq_accumulate<DIM>(I, J, p[0] - cx, p[1] - cy, DIM == 3 ? p[2] - cz : 0., w_ghost[f] - wi);
    }
    double gx, gy, gz;
    lsq_solve<DIM>(I, J, gx, gy, gz);
    wx[i] = gx;
    wy[i] = gy;
    wz[i] = gz;
}

extern "C" int mfv_cell_gradient(const mfv_mesh *m, const double *w_c, const double *w_ghost, double *w_x,
                                 double *w_y, double *w_z, void *stream)
{
    if (int rc = validate_mesh(m, 1, 0, 0)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(w_c && w_ghost && w_x && w_y && w_z, "NULL array");
    auto k = (m->dim == 2) ? k_cell_gradient<2> : k_cell_gradient<3>;
    hipLaunchKernelGGL(k, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), to_kernel(m), w_c,
                       w_ghost, w_x, w_y, w_z);
    return check_launch("k_cell_gradient");
}

// ---------------------------------------------------------------------------------------------
// Barth-Jespersen limiter (ast/functions2d.py:280-325, functions3d.py:364-408)

template <int DIM>
__device__ __forceinline__ double barth_psi(const Mesh &m, int i, double wi, double gx, double gy, double gz,
                                            double w_max, double w_min)
{
    const double eps = (DIM == 2) ? 1e-8 : 1e-10;
    const double cx = m.cell_center[3 * i], cy = m.cell_center[3 * i + 1], cz = m.cell_center[3 * i + 2];
    const int nf = m.cell_nfaces[i];
    double psi = 1., psi_ij = 1.;
    for (int j = 0; j < nf; j++) {
        int f = m.cell_faces[(size_t)j * m.C + i];
        const double *fc = m.face_center + 3 * (size_t)f;
        double r1 = fc[0] - cx, r2 = fc[1] - cy;
        double delta2 = (DIM == 2) ? (gx * r1 + gy * r2) : (gx * r1 + gy * r2 + gz * (fc[2] - cz));
        if (fabs(delta2) < eps) {
            psi_ij = 1.;
        } else {
            if (delta2 > 0.) psi_ij = pmin(1., (w_max - wi) / delta2);
            if (delta2 < 0.) psi_ij = pmin(1., (w_min - wi) / delta2);
        }
        psi = pmin(psi, psi_ij);
    }
    return psi;
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_barthlimiter(Mesh m, const double *__restrict__ w_c,
                                                          const double *__restrict__ w_ghost,
                                                          const double *__restrict__ wx,
                                                          const double *__restrict__ wy,
                                                          const double *__restrict__ wz, double *__restrict__ psi)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    const double wi = w_c[i];
    double w_max = wi, w_min = wi;
    const int nf = m.cell_nfaces[i];
    for (int j = 0; j < nf; j++) {
        int f = m.cell_faces[(size_t)j * m.C + i];
        int nm = m.face_name[f];
        double a = w_c[m.face_cells[2 * f]];
        double b = (nm == 0 || nm > 10) ? w_c[m.face_cells[2 * f + 1]] : w_ghost[f];
        w_max = pmax(pmax(w_max, a), b);
        w_min = pmin(pmin(w_min, a), b);
    }
    psi[i] = barth_psi<DIM>(m, i, wi, wx[i], wy[i], DIM == 3 ? wz[i] : 0., w_max, w_min);
}

extern "C" int mfv_barthlimiter(const mfv_mesh *m, const double *w_c, const double *w_ghost, const double *w_x,
                                const double *w_y, const double *w_z, double *psi, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(w_c && w_ghost && w_x && w_y && w_z && psi, "NULL array");
    auto k = (m->dim == 2) ? k_barthlimiter<2> : k_barthlimiter<3>;
    hipLaunchKernelGGL(k, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), to_kernel(m), w_c,
                       w_ghost, w_x, w_y, w_z, psi);
    return check_launch("k_barthlimiter");
}

// ---------------------------------------------------------------------------------------------
// dt reduction helpers

__global__ void k_dt_reset(double *dt) { dt[0] = 1e6; }

__global__ void k_dt_ring_init(double *ring, const double *dt0)
{
    ring[0] = dt0[0];
    ring[1] = 1e6;
    ring[2] = 1e6;
    ring[3] = 0.;
}

extern "C" int mfv_dt_ring_init(double *dt_ring_p, const double *dt0, void *stream)
{
    MFV_REQUIRE(dt_ring_p && dt0, "NULL array");
    hipLaunchKernelGGL(k_dt_ring_init, dim3(1), dim3(1), 0, as_stream(stream), dt_ring_p, dt0);
    return check_launch("k_dt_ring_init");
}

extern "C" int mfv_dt_reset(double *dt, void *stream)
{
    MFV_REQUIRE(dt, "dt is NULL");
    hipLaunchKernelGGL(k_dt_reset, dim3(1), dim3(1), 0, as_stream(stream), dt);
    return check_launch("k_dt_reset");
}

namespace mfv {

// Per-wave dt partial minima live in one buffer per (device, stream): work queued on one
// stream runs in order, so the buffer is reused without a race, and calls on different
// streams never share one. A buffer only grows. A smaller buffer it replaces is never freed:
// a graph captured on that stream (by the caller, e.g. torch) may still hold its address and
// be replayed later, so retired buffers stay allocated until the process ends. Growth is by
// doubling, so the retired buffers together are smaller than the live one.
struct DtScratch {
    int dev;
    hipStream_t stream;
    double *ptr;
    size_t cap;
};
static std::mutex g_dt_mu;
static std::vector<DtScratch> g_dt_scr;
static std::vector<double *> g_dt_retired;

void graph_free(MfvGraph *g)
{
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    if (g->part) (void)hipFree(g->part);
    delete g;
}

int graph_begin(MfvGraph *g)
{
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        set_error("graph capture setup: %s", hipGetErrorString(e));
        return MFV_EHIP;
    }
    return MFV_OK;
}

int graph_finish(MfvGraph *g, int rc)
{
    hipError_t e = hipStreamEndCapture(g->stream, &g->graph);
    if (rc != MFV_OK) return rc;
    if (e != hipSuccess) {
        set_error("graph capture: %s", hipGetErrorString(e));
        return MFV_EHIP;
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        set_error("graph instantiate: %s", hipGetErrorString(e));
        return MFV_EHIP;
    }
    return MFV_OK;
}

double *dt_scratch(size_t nwaves, hipStream_t st)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("dt scratch: no current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_dt_mu);
    DtScratch *e = nullptr;
    for (auto &x : g_dt_scr)
        if (x.dev == dev && x.stream == st) e = &x;
    if (!e) {
        g_dt_scr.push_back(DtScratch{dev, st, nullptr, 0});
        e = &g_dt_scr.back();
    }
    if (e->cap < nwaves) {
        const size_t cap = e->cap * 2 > nwaves ? e->cap * 2 : nwaves;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
            set_error("dt scratch: the stream is being captured and its partials buffer must grow to %zu doubles; "
                      "run one uncaptured call of the same size first",
                      nwaves);
            return nullptr;
        }
        if (e->ptr) g_dt_retired.push_back(e->ptr);
        e->ptr = nullptr;
        e->cap = 0;
        if (hipMalloc(reinterpret_cast<void **>(&e->ptr), cap * sizeof(double)) != hipSuccess) {
            set_error("dt scratch: hipMalloc of %zu doubles failed", cap);
            e->ptr = nullptr;
            return nullptr;
        }
        e->cap = cap;
    }
    return e->ptr;
}

__global__ void __launch_bounds__(1024) k_dt_finish(const double *__restrict__ part, size_t n, double *__restrict__ out)
{
    __shared__ double sh[1024];
    // four partials in flight per thread (the minimum of a set does not depend on the order)
    double m0 = __longlong_as_double(0x7ff0000000000000LL), m1 = m0, m2 = m0, m3 = m0;
    size_t j = threadIdx.x;
    for (; j + 3 * 1024 < n; j += 4 * 1024) {
        const double a = part[j], b = part[j + 1024], c = part[j + 2048], e = part[j + 3072];
        m0 = (a < m0) ? a : m0;
        m1 = (b < m1) ? b : m1;
        m2 = (c < m2) ? c : m2;
        m3 = (e < m3) ? e : m3;
    }
    for (; j < n; j += 1024) {
        const double v = part[j];
        m0 = (v < m0) ? v : m0;
    }
    m0 = (m1 < m0) ? m1 : m0;
    m2 = (m3 < m2) ? m3 : m2;
    const double m = (m2 < m0) ? m2 : m0;
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double o = sh[threadIdx.x + w];
            sh[threadIdx.x] = (o < sh[threadIdx.x]) ? o : sh[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double r = sh[0];
        if (r < out[0]) out[0] = r;
    }
}

int dt_finish(const double *part, size_t n, double *out, hipStream_t st)
{
    if (int rc = check_launch("dt kernel")) return rc;
    hipLaunchKernelGGL(k_dt_finish, dim3(1), dim3(1024), 0, st, part, n, out);
    return check_launch("k_dt_finish");
}

}  // namespace mfv

int mfv::finish_dt(double *dt_dev, double *dt_host, void *stream)
{
    if (int rc = check_launch("dt kernel")) return rc;
    if (dt_host) {
        hipError_t e = hipMemcpyAsync(dt_host, dt_dev, sizeof(double), hipMemcpyDeviceToHost, as_stream(stream));
        if (e == hipSuccess) e = hipStreamSynchronize(as_stream(stream));
        if (e != hipSuccess) {
            set_error("dt readback: %s", hipGetErrorString(e));
            return MFV_EHIP;
        }
    }
    return MFV_OK;
}

// ---------------------------------------------------------------------------------------------
// advection (solvers/advec/fvm_utils.py)

// time_step (fvm_utils.py:293-309): the per-cell candidate; faces summed in faceid order.
__device__ __forceinline__ double adv_cell_dt(const Mesh &m, int i, double u, double v, double w, double cfl)
{
    const int nf = m.cell_nfaces[i];
    double lam = 0.;
    for (int j = 0; j < nf; j++) {
        const double *n = m.face_normal + 3 * (size_t)m.cell_faces[(size_t)j * m.C + i];
        lam += fabs(u * n[0] + v * n[1] + w * n[2]);
    }
    return cfl * m.cell_volume[i] / lam;
}

__global__ void __launch_bounds__(kBlock) k_adv_time_step(Mesh m, const double *__restrict__ u,
                                                           const double *__restrict__ v,
                                                           const double *__restrict__ w, double cfl,
                                                           double *__restrict__ dtpart)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = (i < m.C) ? adv_cell_dt(m, i, u[i], v[i], w[i], cfl) : 1e300;
    wave_min_part(dtpart, (size_t)i, d);
}

extern "C" int mfv_time_step(const mfv_mesh *m, const double *u, const double *v, const double *w, double cfl,
                             double *dt_dev, double *dt_host, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(u && v && w && dt_dev, "NULL array");
    if (int rc = mfv_dt_reset(dt_dev, stream)) return rc;
    if (m->nb_cells) {
        const size_t nw = waves_for(m->nb_cells);
        double *part = dt_scratch(nw, as_stream(stream));
        if (!part) return MFV_ENOMEM;
        hipLaunchKernelGGL(k_adv_time_step, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                           to_kernel(m), u, v, w, cfl, part);
        if (int rc = dt_finish(part, nw, dt_dev, as_stream(stream))) return rc;
    }
    return finish_dt(dt_dev, dt_host, stream);
}

// Upwind flux of face f in its own orientation (compute_upwind_flux, fvm_utils.py:5-17) with
// the MUSCL reconstruction of explicitscheme_convective_{2d,3d}. Left = face_cells[f][0];
// the right value comes from face_cells[f][1] or, on a boundary face, from w_ghost[f].
// ORDER 1 skips the gradient reads: the reference's term (order-1)*psi*(grad.r) is then
// 0 * 0 * 0 (psi and gradients stay zero when compute_cell_gradient is never called).
// OWNER_GHOST takes the Neumann ghost w_ghost[f] = w_c[face_cells[f][0]] on the fly.
template <int DIM, int ORDER, bool OWNER_GHOST>
__device__ __forceinline__ double adv_face_flux(const Mesh &m, int f, const double *__restrict__ w_c,
                                                const double *__restrict__ w_ghost, double uf, double vf,
                                                double wf, const double *__restrict__ gx,
                                                const double *__restrict__ gy, const double *__restrict__ gz,
                                                const double *__restrict__ psi, double om, int &a)
{
    a = m.face_cells[2 * f];
    const int b = m.face_cells[2 * f + 1];
    double w_l = w_c[a];
    double w_r = (b >= 0) ? w_c[b] : (OWNER_GHOST ? w_c[a] : w_ghost[f]);
    if (ORDER == 1) {
        w_l = w_l + 0.;
        if (b >= 0) w_r = w_r + 0.;
    } else {
        const double *fc = m.face_center + 3 * (size_t)f;
        const double *ca = m.cell_center + 3 * (size_t)a;
        double rl0 = fc[0] - ca[0], rl1 = fc[1] - ca[1];
        if (DIM == 2)
            w_l = w_l + om * psi[a] * (gx[a] * rl0 + gy[a] * rl1);
        else
            w_l = w_l + om * psi[a] * (gx[a] * rl0 + gy[a] * rl1 + gz[a] * (fc[2] - ca[2]));
        if (b >= 0) {
            const double *cb = m.cell_center + 3 * (size_t)b;
            double rr0 = fc[0] - cb[0], rr1 = fc[1] - cb[1];
            if (DIM == 2)
                w_r = w_r + om * psi[b] * (gx[b] * rr0 + gy[b] * rr1);
            else
                w_r = w_r + om * psi[b] * (gx[b] * rr0 + gy[b] * rr1 + gz[b] * (fc[2] - cb[2]));
        }
    }
    const double *n = m.face_normal + 3 * (size_t)f;
    double sign = uf * n[0] + vf * n[1] + wf * n[2];
    double sol = (sign >= 0) ? w_l : w_r;
    return sign * sol;
}

// Residual of cell i: contributions added in the reference's face order (inner faces by
// increasing id, then boundary faces by increasing id) so the sum is bit-identical to the
// face-ordered scatter rez[c0] -= F; rez[c1] += F.
template <int DIM, int ORDER, bool OWNER_GHOST>
__device__ __forceinline__ double adv_residual(const Mesh &m, int i, const double *__restrict__ w_c,
                                               const double *__restrict__ w_ghost, const double *__restrict__ uf,
                                               const double *__restrict__ vf, const double *__restrict__ wf,
                                               const double *__restrict__ gx, const double *__restrict__ gy,
                                               const double *__restrict__ gz, const double *__restrict__ psi,
                                               double om)
{
    const int nf = m.cell_nfaces[i];
    double rez = 0.;
    for (int r = 0; r < nf; r++) {
        int s = m.order_split[(size_t)r * m.C + i];
        int f = m.cell_faces[(size_t)s * m.C + i];
        int a;
        double flux =
            adv_face_flux<DIM, ORDER, OWNER_GHOST>(m, f, w_c, w_ghost, uf[f], vf[f], wf[f], gx, gy, gz, psi, om, a);
        if (a == i)
            rez -= flux;
        else
            rez += flux;
    }
    return rez;
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_adv_residual(Mesh m, double *__restrict__ rez_w,
                                                          const double *__restrict__ w_c,
                                                          const double *__restrict__ w_ghost,
                                                          const double *__restrict__ uf,
                                                          const double *__restrict__ vf,
                                                          const double *__restrict__ wf,
                                                          const double *__restrict__ gx,
                                                          const double *__restrict__ gy,
                                                          const double *__restrict__ gz,
                                                          const double *__restrict__ psi, double om)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    // the kernel-level entry point always evaluates the reference expression with its arrays
    rez_w[i] = adv_residual<DIM, 2, false>(m, i, w_c, w_ghost, uf, vf, wf, gx, gy, gz, psi, om);
}

extern "C" int mfv_explicitscheme_convective(const mfv_mesh *m, double *rez_w, const double *w_c,
                                             const double *w_ghost, const double *u_face, const double *v_face,
                                             const double *w_face, const double *w_x, const double *w_y,
                                             const double *w_z, const double *psi, int order, void *stream)
{
    if (int rc = validate_mesh(m, 0, 1, 0)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(rez_w && w_c && w_ghost && u_face && v_face && w_face && w_x && w_y && w_z && psi,
                "NULL array");
    double om = (double)(order - 1);
    auto k = (m->dim == 2) ? k_adv_residual<2> : k_adv_residual<3>;
    hipLaunchKernelGGL(k, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), to_kernel(m), rez_w,
                       w_c, w_ghost, u_face, v_face, w_face, w_x, w_y, w_z, psi, om);
    return check_launch("k_adv_residual");
}

__global__ void __launch_bounds__(kBlock) k_update_new_value(int C, double *__restrict__ ne_c,
                                                              const double *__restrict__ rez,
                                                              const double *__restrict__ dissip,
                                                              const double *__restrict__ src, double dtime,
                                                              const double *__restrict__ dt_dev,
                                                              const double *__restrict__ vol)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    double dt = dt_dev ? dt_dev[0] : dtime;
    ne_c[i] += dt * ((rez[i] + dissip[i]) / vol[i] + src[i]);
}

extern "C" int mfv_update_new_value(const mfv_mesh *m, double *ne_c, const double *rez, const double *dissip,
                                    const double *src, double dtime, const double *dt_dev, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(ne_c && rez && dissip && src, "NULL array");
    hipLaunchKernelGGL(k_update_new_value, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       m->nb_cells, ne_c, rez, dissip, src, dtime, dt_dev, m->cell_volume);
    return check_launch("k_update_new_value");
}

// ----- fused advection step ------------------------------------------------------------------
// Neumann ghosts (w_ghost[f] = w[c0]) are read straight from the owner cell.

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_adv_grad_limit(Mesh m, const double *__restrict__ w,
                                                            double *__restrict__ scratch)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    const double cx = m.cell_center[3 * i], cy = m.cell_center[3 * i + 1], cz = m.cell_center[3 * i + 2];
    const double wi = w[i];
    double I[6] = {0., 0., 0., 0., 0., 0.}, J[3] = {0., 0., 0.};
    const int nn = m.lsq_ncells[i];
    for (int j = 0; j < nn; j++) {
        int c = m.lsq_cells[(size_t)j * m.C + i];
        const double *p = m.cell_center + 3 * (size_t)c;
        lsq_accumulate<DIM>(I, J, p[0] - cx, p[1] - cy, DIM == 3 ? p[2] - cz : 0., w[c] - wi);
    }
    const int ng = m.lsq_nghosts[i];
    for (int j = 0; j < ng; j++) {
        int f = m.lsq_ghosts[(size_t)j * m.C + i];
        const double *p = m.face_ghost + 3 * (size_t)f;
        lsq_accumulate<DIM>(I, J, p[0] - cx, p[1] - cy, DIM == 3 ? p[2] - cz : 0., w[m.face_cells[2 * f]] - wi);
    }
    double gx, gy, gz;
    lsq_solve<DIM>(I, J, gx, gy, gz);
    double w_max = wi, w_min = wi;
    const int nf = m.cell_nfaces[i];
    for (int j = 0; j < nf; j++) {
        int f = m.cell_faces[(size_t)j * m.C + i];
        int nm = m.face_name[f];
        double a = w[m.face_cells[2 * f]];
        double b = (nm == 0 || nm > 10) ? w[m.face_cells[2 * f + 1]] : w[m.face_cells[2 * f]];
        w_max = pmax(pmax(w_max, a), b);
        w_min = pmin(pmin(w_min, a), b);
    }
    scratch[i] = gx;
    scratch[(size_t)m.C + i] = gy;
    if (DIM == 3) scratch[2 * (size_t)m.C + i] = gz;
    scratch[(size_t)DIM * m.C + i] = barth_psi<DIM>(m, i, wi, gx, gy, gz, w_max, w_min);
}

template <int DIM, int ORDER>
__global__ void __launch_bounds__(kBlock) k_adv_step(Mesh m, const double *__restrict__ w_in,
                                                      double *__restrict__ w_out, const double *__restrict__ uc,
                                                      const double *__restrict__ vc,
                                                      const double *__restrict__ wc,
                                                      const double *__restrict__ uf,
                                                      const double *__restrict__ vf,
                                                      const double *__restrict__ wf, double cfl, double om,
                                                      const double *__restrict__ scratch,
                                                      DtRing ring, double *__restrict__ dtpart)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    ring_advance(ring);
    double dloc = 1e300;
    if (i < m.C) {
        const double *gx = scratch, *gy = scratch + m.C, *gz = scratch + 2 * (size_t)m.C;
        const double *psi = scratch + (size_t)DIM * m.C;
        double rez = adv_residual<DIM, ORDER, true>(m, i, w_in, nullptr, uf, vf, wf, gx, gy, gz, psi, om);
        double dt = ring.cur[0];
        // update_new_value with the dissipative and source terms zero
        w_out[i] = w_in[i] + dt * ((rez + 0.) / m.cell_volume[i] + 0.);
        dloc = adv_cell_dt(m, i, uc[i], vc[i], wc[i], cfl);
    }
    wave_min_part(dtpart, (size_t)i, dloc);
}

extern "C" int mfv_advection_step(const mfv_mesh *m, const double *w_in, double *w_out, const double *u_c,
                                  const double *v_c, const double *w_c, const double *u_face,
                                  const double *v_face, const double *w_face, double cfl, int order,
                                  double *dt_ring_p, int step, double *scratch, void *stream)
{
    if (int rc = validate_mesh(m, order == 2, 1, 0)) return rc;
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    MFV_REQUIRE(w_in && w_out && w_in != w_out && u_c && v_c && w_c && u_face && v_face && w_face && dt_ring_p,
                "NULL or aliased array");
    DtRing r = dt_ring(dt_ring_p, step);
    MFV_REQUIRE(order == 1 || scratch, "order 2 needs a scratch of (dim+1)*C doubles");
    MFV_REQUIRE(!(m->dim == 2 && m->max_cell_faces > 4), "2D cells have at most 4 faces");
    if (m->nb_cells == 0) return MFV_OK;
    Mesh k = to_kernel(m);
    hipStream_t st = as_stream(stream);
    if (order == 2) {
        auto kg = (m->dim == 2) ? k_adv_grad_limit<2> : k_adv_grad_limit<3>;
        hipLaunchKernelGGL(kg, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, st, k, w_in, scratch);
        if (int rc = check_launch("k_adv_grad_limit")) return rc;
    }
    const size_t nw = waves_for(m->nb_cells);
    double *part = dt_scratch(nw, as_stream(stream));
    if (!part) return MFV_ENOMEM;
    auto ks = (m->dim == 2) ? ((order == 2) ? k_adv_step<2, 2> : k_adv_step<2, 1>)
                            : ((order == 2) ? k_adv_step<3, 2> : k_adv_step<3, 1>);
    hipLaunchKernelGGL(ks, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, st, k, w_in, w_out, u_c, v_c, w_c,
                       u_face, v_face, w_face, cfl, (double)(order - 1), scratch, r, part);
    if (int rc = check_launch("k_adv_step")) return rc;
    return dt_finish(part, nw, r.next, st);
}
