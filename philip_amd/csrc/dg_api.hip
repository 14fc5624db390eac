// C ABI of the device DG path (declared in include/philip_dg.h) and the host
// set-up it needs: operators, connectivity, buffers, RK register management.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <chrono>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "dg_host.hpp"

namespace dgx {

static void legendre_pd(int n, double x, double *pn, double *dpn)
{
    double p0 = 1.0, p1 = x;
    if (n == 0) { *pn = 1.0; *dpn = 0.0; return; }
    for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    *pn = p1;
    *dpn = (std::fabs(x) < 1.0) ? n * (x * p1 - p0) / (x * x - 1.0) : 0.5 * n * (n + 1.0);
}

void gauss_lobatto(int n, double *x, double *w)
{
    const int N = n - 1;
    double t[64];
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * i / N);
        if (i > 0 && i < N) {
            for (int it = 0; it < 100; ++it) {
                double pn, dpn;
                legendre_pd(N, z, &pn, &dpn);
                const double d2 = (2.0 * z * dpn - N * (N + 1.0) * pn) / (1.0 - z * z);
                const double dz = dpn / d2;
                z -= dz;
                if (std::fabs(dz) < 1e-17) break;
            }
        }
        t[i] = z;
    }
    t[0] = 1.0;
    t[N] = -1.0;
    for (int i = 0; i < n; ++i) {
        double pn, dpn;
        legendre_pd(N, t[i], &pn, &dpn);
        x[i] = 0.5 * (1.0 - t[i]);
        w[i] = 1.0 / (N * (N + 1.0) * pn * pn);
    }
    for (int i = 0; i < n / 2; ++i) {
        const double a = 0.5 * (x[i] + 1.0 - x[n - 1 - i]);
        x[i] = a;
        x[n - 1 - i] = 1.0 - a;
        const double b = 0.5 * (w[i] + w[n - 1 - i]);
        w[i] = b;
        w[n - 1 - i] = b;
    }
    x[0] = 0.0;
    x[n - 1] = 1.0;
    if (n % 2 == 1) x[n / 2] = 0.5;
}

void gauss_legendre(int n, double *x, double *w)
{
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double pn, dpn;
            legendre_pd(n, z, &pn, &dpn);
            const double dz = pn / dpn;
            z -= dz;
            if (std::fabs(dz) < 1e-17) break;
        }
        double pn, dpn;
        legendre_pd(n, z, &pn, &dpn);
        x[i] = 0.5 * (1.0 - z);
        w[i] = 1.0 / ((1.0 - z * z) * dpn * dpn);
    }
    for (int i = 0; i < n / 2; ++i) {
        const double a = 0.5 * (x[i] + 1.0 - x[n - 1 - i]);
        x[i] = a;
        x[n - 1 - i] = 1.0 - a;
        const double b = 0.5 * (w[i] + w[n - 1 - i]);
        w[i] = b;
        w[n - 1 - i] = b;
    }
    if (n % 2 == 1) x[n / 2] = 0.5;
}

void lagrange(int n, const double *s, double x, double *val, double *der)
{
    for (int j = 0; j < n; ++j) {
        double v = 1.0;
        for (int m = 0; m < n; ++m)
            if (m != j) v *= (x - s[m]) / (s[j] - s[m]);
        double d = 0.0;
        for (int k = 0; k < n; ++k) {
            if (k == j) continue;
            double t = 1.0 / (s[j] - s[k]);
            for (int m = 0; m < n; ++m)
                if (m != j && m != k) t *= (x - s[m]) / (s[j] - s[m]);
            d += t;
        }
        if (val) val[j] = v;
        if (der) der[j] = d;
    }
}

Ops1D build_ops(int n)
{
    Ops1D o{};
    double x[MAXN1D], w[MAXN1D];
    gauss_lobatto(n, x, w);
    for (int a = 0; a < n; ++a) {
        double val[MAXN1D], der[MAXN1D];
        lagrange(n, x, x[a], val, der);
        for (int b = 0; b < n; ++b) o.D[a][b] = der[b];
        o.w[a] = w[a];
    }
    // flux-basis stiffness Q = W D (operators.cpp:1440-1477), skew form scaled by 1/w_a
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) o.Sk[a][b] = (w[a] * o.D[a][b] - w[b] * o.D[b][a]) / w[a];
    o.inv_wb = 1.0 / w[0];
    for (int a = 0; a < MAXN1D; ++a)
        for (int b = 0; b < MAXN1D; ++b) {
            o.Skf[a][b] = (float)o.Sk[a][b];
            o.Df[a][b] = (float)o.D[a][b];
        }
    o.inv_wbf = (float)o.inv_wb;
    return o;
}

int Arena::init(int64_t bytes)
{
    bytes = (bytes + 255) & ~int64_t(255);
    void *p = nullptr;
    if (hipMalloc(&p, (size_t)bytes) != hipSuccess) return DGX_ERR_HIP;
    base_ = static_cast<char *>(p);
    size_ = bytes;
    free_.clear();
    used_.clear();
    free_[0] = bytes;
    used_bytes_ = 0;
    return DGX_OK;
}

void Arena::release()
{
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
    size_ = used_bytes_ = 0;
    free_.clear();
    used_.clear();
}

void *Arena::alloc(int64_t bytes)
{
    bytes = (bytes + 255) & ~int64_t(255);
    if (bytes <= 0) bytes = 256;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        const int64_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        used_[off] = bytes;
        used_bytes_ += bytes;
        return base_ + off;
    }
    return nullptr;
}

void Arena::free(void *p)
{
    if (!p) return;
    const int64_t off = static_cast<char *>(p) - base_;
    auto u = used_.find(off);
    if (u == used_.end()) return;
    int64_t start = off, len = u->second;
    used_bytes_ -= len;
    used_.erase(u);
    auto next = free_.lower_bound(start);
    if (next != free_.end() && next->first == start + len) {  // merge with the following range
        len += next->second;
        next = free_.erase(next);
    }
    if (next != free_.begin()) {  // merge with the preceding range
        auto prev = std::prev(next);
        if (prev->first + prev->second == start) {
            start = prev->first;
            len += prev->second;
            free_.erase(prev);
        }
    }
    free_[start] = len;
}

} // namespace dgx

using namespace dgx;

// device buffers of a context: from its arena (serving) or hipMalloc
static void *dev_alloc(dgx_ctx *c, size_t bytes)
{
    if (c->arena) return c->arena->alloc((int64_t)bytes);
    void *p = nullptr;
    return hipMalloc(&p, bytes) == hipSuccess ? p : nullptr;
}

static void dev_free(dgx_ctx *c, void *p)
{
    if (!p) return;
    if (c->arena) c->arena->free(p);
    else (void)hipFree(p);
}

#define HIPCHK(call)                                                                     \
    do {                                                                                 \
        hipError_t _e = (call);                                                          \
        if (_e != hipSuccess) {                                                          \
            c->err = std::string(#call) + ": " + hipGetErrorString(_e);                   \
            return DGX_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)

static int fail(dgx_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

extern "C" {

void dgx_default_params(dgx_params *p)
{
    std::memset(p, 0, sizeof(*p));
    p->dim = 3;
    p->poly_degree = 2;
    p->flux_nodes = DGX_GLL;
    p->use_split_form = 0;
    p->two_point_flux = DGX_KG;
    p->conv_flux = DGX_LAX_FRIEDRICHS;
    p->pde = DGX_EULER;
    p->diss_flux = DGX_SIPG;
    p->use_constant_viscosity = 0;
    p->gamma_gas = 1.4;
    p->mach_inf = 0.1;
    p->reynolds_inf = 1600.0;
    p->prandtl = 0.72;
    p->constant_viscosity = 1.0;
    p->temperature_inf = 273.15;
    p->sipg_penalty_factor = 1.0;
    p->n_cells_1d = 4;
    p->domain_left = 0.0;
    p->domain_right = 6.28318530717958623200;
    p->cell_order = DGX_ORDER_DEALII;
    p->precision = DGX_FP64;
    for (int f = 0; f < 6; ++f) p->boundary_type[f] = 0;
    p->angle_of_attack = 0.0;
    p->grid_warp = 0;
    p->thermal_bc = 0;
    p->wall_temperature = 1.0;
}

// DGFactory::create_discontinuous_galerkin + allocate_system: operators, device buffers (from
// `arena` when given), connectivity computed on the device; all work enqueued on `stream`.
static int create_ctx(const dgx_params *p, dgx::Arena *arena, hipStream_t stream, dgx_ctx **out)
{
    if (!p || !out) return DGX_ERR_INVALID;
    *out = nullptr;
    auto bad = [&](int code, const std::string &m) {
        std::fprintf(stderr, "dgx_create: %s\n", m.c_str());
        return code;
    };
    if (p->dim != 2 && p->dim != 3) return bad(DGX_ERR_INVALID, "dim must be 2 or 3");
    if (p->poly_degree < 1 || p->poly_degree > 6) return bad(DGX_ERR_UNSUPPORTED, "device path supports p = 1..6");
    if (p->flux_nodes != DGX_GLL) return bad(DGX_ERR_UNSUPPORTED, "device path supports collocated GLL flux nodes only");
    if (p->n_cells_1d < 2) return bad(DGX_ERR_INVALID, "n_cells_1d must be >= 2");
    if (p->conv_flux < 0 || p->conv_flux > 5) return bad(DGX_ERR_INVALID, "unknown conv_flux");
    if (p->conv_flux >= 3 && !p->use_split_form)
        return bad(DGX_ERR_INVALID, "two point flux and not using split form are not compatible");
    if (p->pde == DGX_NAVIER_STOKES && p->dim != 3) return bad(DGX_ERR_UNSUPPORTED, "Navier-Stokes device path is 3D");
    if (p->precision != DGX_FP64 && p->precision != DGX_MIXED) return bad(DGX_ERR_INVALID, "unknown precision");
    for (int f = 0; f < 2 * p->dim; ++f) {
        const int t = p->boundary_type[f], o = p->boundary_type[f ^ 1];
        if ((t == 0) != (o == 0)) return bad(DGX_ERR_INVALID, "both faces of a direction periodic or both boundaries");
        if (t && (t < 1001 || t > 1007)) return bad(DGX_ERR_UNSUPPORTED, "boundary types 1001..1007 only");
        if (t && p->pde != DGX_EULER && t != 1001 && t != 1006)
            return bad(DGX_ERR_UNSUPPORTED, "Navier-Stokes boundaries: the walls 1001 and 1006 only");
    }
    if (p->grid_warp != 0 && p->grid_warp != 1) return bad(DGX_ERR_INVALID, "grid_warp is 0 or 1");
    if (p->thermal_bc != 0 && p->thermal_bc != 1) return bad(DGX_ERR_INVALID, "thermal_bc is 0 or 1");
    if (!(p->wall_temperature > 0.0)) return bad(DGX_ERR_INVALID, "wall_temperature must be positive");
    dgx_ctx *c = new dgx_ctx();
    c->P = *p;
    c->arena = arena;
    c->stream = stream;
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) {
        delete c;
        return bad(DGX_ERR_HIP, "no HIP device");
    }
    c->dim = p->dim;
    c->p = p->poly_degree;
    c->n = c->p + 1;
    c->Np = (c->dim == 3) ? c->n * c->n * c->n : c->n * c->n;
    c->ns = c->dim + 2;
    const int N = p->n_cells_1d;
    c->ncells = (c->dim == 3) ? N * N * N : N * N;
    c->ndofs = (int64_t)c->ncells * c->ns * c->Np;
    c->h = (p->domain_right - p->domain_left) / N;
    c->ops = build_ops(c->n);
    PhysConst &pc = c->pc;
    pc.gam = p->gamma_gas;
    pc.gamm1 = p->gamma_gas - 1.0;
    pc.mach2 = p->mach_inf * p->mach_inf;
    pc.reynolds = p->reynolds_inf;
    pc.prandtl = p->prandtl;
    pc.mu_const = p->constant_viscosity;
    pc.temperature_ratio = 110.4 / p->temperature_inf;
    pc.use_constant_viscosity = p->use_constant_viscosity;
    pc.two_point = p->two_point_flux;
    pc.conv_flux = p->conv_flux;
    pc.diss_flux = p->diss_flux;
    pc.vinf[0] = std::cos(p->angle_of_attack);
    pc.vinf[1] = std::sin(p->angle_of_attack);
    pc.vinf[2] = 0.0;
    pc.thermal_bc = p->thermal_bc;
    pc.wall_temperature = p->wall_temperature;
    for (int f = 0; f < 2 * c->dim; ++f) c->has_boundary = c->has_boundary || p->boundary_type[f] != 0;
    c->nbr_count = (int64_t)c->ncells * 2 * c->dim;
    const size_t bytes = sizeof(double) * (size_t)c->ndofs;
    c->nbr_dev = static_cast<int *>(dev_alloc(c, sizeof(int) * (size_t)c->nbr_count));
    c->u_dev = static_cast<double *>(dev_alloc(c, bytes));
    c->r_dev = static_cast<double *>(dev_alloc(c, bytes));
    c->status_dev = static_cast<int *>(dev_alloc(c, 256));
    bool ok = c->nbr_dev && c->u_dev && c->r_dev && c->status_dev;
    if (ok && p->pde == DGX_NAVIER_STOKES)
        for (int k = 0; k < c->dim; ++k) {
            if (p->precision == DGX_MIXED) c->auxf_dev[k] = static_cast<float *>(dev_alloc(c, sizeof(float) * (size_t)c->ndofs));
            else c->aux_dev[k] = static_cast<double *>(dev_alloc(c, bytes));
            ok = ok && (c->auxf_dev[k] || c->aux_dev[k]);
        }
    if (!ok) {
        c->stream = nullptr;
        dgx_destroy(c);
        return arena ? DGX_ERR_STATE : bad(DGX_ERR_HIP, "hipMalloc failed");  // arena full: try later
    }
    if (hipMemsetAsync(c->status_dev, 0, 256, stream) != hipSuccess ||
        hipMemsetAsync(c->u_dev, 0, bytes, stream) != hipSuccess ||
        hipMemsetAsync(c->r_dev, 0, bytes, stream) != hipSuccess || launch_mesh(c) != DGX_OK) {
        dgx_destroy(c);
        return bad(DGX_ERR_HIP, "device set-up failed");
    }
    if (p->grid_warp) {
        // the reference's curved periodic grid: node coordinates and metric terms on the device
        c->xn_dev = static_cast<double *>(dev_alloc(c, sizeof(double) * 3 * (size_t)c->ncells * c->Np));
        c->met_dev = static_cast<double *>(
            dev_alloc(c, sizeof(double) * (size_t)(c->dim * c->dim + 1) * c->ncells * c->Np));
        c->pen_dev = static_cast<double *>(dev_alloc(c, sizeof(double) * (size_t)c->dim * c->ncells));
        if (!c->xn_dev || !c->met_dev || !c->pen_dev || launch_coords(c, c->xn_dev, 1) != DGX_OK || launch_metrics(c) != DGX_OK) {
            c->stream = nullptr;
            dgx_destroy(c);
            return arena ? DGX_ERR_STATE : bad(DGX_ERR_HIP, "curvilinear grid set-up failed");
        }
        c->curv = true;
    }
    *out = c;
    return DGX_OK;
}

int dgx_create(const dgx_params *p, dgx_ctx **out)
{
    dgx_ctx *c = nullptr;
    int rc = create_ctx(p, nullptr, nullptr, &c);
    if (rc) return rc;
    if (hipStreamSynchronize(nullptr) != hipSuccess) {
        dgx_destroy(c);
        return DGX_ERR_HIP;
    }
    *out = c;
    return DGX_OK;
}

int dgx_destroy(dgx_ctx *c)
{
    if (!c) return DGX_OK;
    (void)hipStreamSynchronize(c->stream);  // teardown: errors are not actionable here
    dev_free(c, c->nbr_dev);
    dev_free(c, c->u_dev);
    dev_free(c, c->r_dev);
    dev_free(c, c->status_dev);
    dev_free(c, c->xn_dev);
    dev_free(c, c->met_dev);
    dev_free(c, c->pen_dev);
    for (int k = 0; k < 3; ++k) {
        dev_free(c, c->aux_dev[k]);
        dev_free(c, c->auxf_dev[k]);
    }
    for (double *r : c->regs) dev_free(c, r);
    (void)hipFree(c->minv_dev);
    (void)hipFree(c->iq_ops);
    (void)hipFree(c->iq_partial);
    (void)hipFree(c->vr_partial);
    (void)hipFree(c->rrk_partial);
    if (c->timing.a) (void)hipEventDestroy(c->timing.a);
    if (c->timing.b) (void)hipEventDestroy(c->timing.b);
    delete c;
    return DGX_OK;
}

const char *dgx_last_error(const dgx_ctx *c) { return c ? c->err.c_str() : "null context"; }
int64_t dgx_n_dofs(const dgx_ctx *c) { return c ? c->ndofs : 0; }
int dgx_n_cells(const dgx_ctx *c) { return c ? c->ncells : 0; }

int dgx_set_solution(dgx_ctx *c, const double *host_u)
{
    if (!c || !host_u) return DGX_ERR_INVALID;
    HIPCHK(hipMemcpyAsync(c->u_dev, host_u, sizeof(double) * c->ndofs, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DGX_OK;
}

int dgx_get_solution(const dgx_ctx *cc, double *host_u)
{
    dgx_ctx *c = const_cast<dgx_ctx *>(cc);
    if (!c || !host_u) return DGX_ERR_INVALID;
    HIPCHK(hipMemcpyAsync(host_u, c->u_dev, sizeof(double) * c->ndofs, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DGX_OK;
}

int dgx_get_right_hand_side(const dgx_ctx *cc, double *host_r)
{
    dgx_ctx *c = const_cast<dgx_ctx *>(cc);
    if (!c || !host_r) return DGX_ERR_INVALID;
    HIPCHK(hipMemcpyAsync(host_r, c->r_dev, sizeof(double) * c->ndofs, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DGX_OK;
}

int dgx_get_auxiliary_solution(const dgx_ctx *cc, int idim, double *host_q)
{
    dgx_ctx *c = const_cast<dgx_ctx *>(cc);
    if (!c || !host_q || idim < 0 || idim >= c->dim) return DGX_ERR_INVALID;
    if (!c->aux_dev[idim] && !c->auxf_dev[idim]) return fail(c, DGX_ERR_STATE, "no auxiliary solution (Euler)");
    if (c->aux_dev[idim]) {
        HIPCHK(hipMemcpyAsync(host_q, c->aux_dev[idim], sizeof(double) * c->ndofs, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } else {  // mixed path: q is held in float
        std::vector<float> qf((size_t)c->ndofs);
        HIPCHK(hipMemcpyAsync(qf.data(), c->auxf_dev[idim], sizeof(float) * c->ndofs, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (size_t k = 0; k < qf.size(); ++k) host_q[k] = qf[k];
    }
    return DGX_OK;
}

int dgx_get_neighbors(const dgx_ctx *cc, int32_t *host_nbr)
{
    dgx_ctx *c = const_cast<dgx_ctx *>(cc);
    if (!c || !host_nbr) return DGX_ERR_INVALID;
    // copy back what the kernels read, so integer parity is checked on the device copy
    HIPCHK(hipMemcpyAsync(host_nbr, c->nbr_dev, sizeof(int) * c->nbr_count, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DGX_OK;
}

int dgx_set_grid_nodes(dgx_ctx *c, const double *host_xyz)
{
    if (!c || !host_xyz) return DGX_ERR_INVALID;
    if (c->batch_locked) {
        c->err = "the context is captured in a batch graph";
        return DGX_ERR_STATE;
    }
    const size_t nx = 3 * (size_t)c->ncells * c->Np, nm = (size_t)(c->dim * c->dim + 1) * c->ncells * c->Np;
    if (!c->xn_dev) c->xn_dev = static_cast<double *>(dev_alloc(c, sizeof(double) * nx));
    if (!c->met_dev) c->met_dev = static_cast<double *>(dev_alloc(c, sizeof(double) * nm));
    if (!c->pen_dev) c->pen_dev = static_cast<double *>(dev_alloc(c, sizeof(double) * (size_t)c->dim * c->ncells));
    if (!c->xn_dev || !c->met_dev || !c->pen_dev) {
        c->err = "device allocation failed";
        return DGX_ERR_HIP;
    }
    HIPCHK(hipMemcpyAsync(c->xn_dev, host_xyz, sizeof(double) * nx, hipMemcpyHostToDevice, c->stream));
    const int rc = launch_metrics(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->curv = true;
    return DGX_OK;
}

int dgx_get_grid_nodes(dgx_ctx *c, double *host_xyz)
{
    if (!c || !host_xyz) return DGX_ERR_INVALID;
    const size_t nx = 3 * (size_t)c->ncells * c->Np;
    if (!c->curv) {  // the affine box: computed into a temporary buffer
        double *tmp = nullptr;
        HIPCHK(hipMalloc(&tmp, sizeof(double) * nx));
        int rc = launch_coords(c, tmp, 0);
        if (!rc && hipMemcpyAsync(host_xyz, tmp, sizeof(double) * nx, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = DGX_ERR_HIP;
        if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = DGX_ERR_HIP;
        (void)hipFree(tmp);
        return rc;
    }
    HIPCHK(hipMemcpyAsync(host_xyz, c->xn_dev, sizeof(double) * nx, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DGX_OK;
}

int dgx_get_metric_terms(dgx_ctx *c, double *host_met)
{
    if (!c || !host_met) return DGX_ERR_INVALID;
    if (!c->curv) {
        c->err = "no curvilinear grid: the metric of the affine box is h^(dim-1) I, det J = h^dim";
        return DGX_ERR_STATE;
    }
    HIPCHK(hipMemcpyAsync(host_met, c->met_dev, sizeof(double) * (size_t)(c->dim * c->dim + 1) * c->ncells * c->Np,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DGX_OK;
}

static StageArgs base_args(dgx_ctx *c, const double *u)
{
    StageArgs a{};
    a.u = u;
    for (int k = 0; k < 3; ++k) {
        a.aux[k] = c->aux_dev[k];
        a.auxf[k] = c->auxf_dev[k];
    }
    a.nbr = c->nbr_dev;
    a.ncells = c->ncells;
    a.inv_h = 1.0 / c->h;
    const unsigned deg = (unsigned)c->p;
    const double pen = (double)(deg * (deg + 1)) / c->h * c->P.sipg_penalty_factor; // dg_base.cpp:397-413
    a.penalty = 0.5 * (pen + pen);
    a.resid_scale = (c->dim == 3) ? c->h * c->h * c->h : c->h * c->h;
    a.status = c->status_dev;
    a.met = c->curv ? c->met_dev : nullptr;
    a.pen = c->curv ? c->pen_dev : nullptr;
    return a;
}

// Reads and clears the non-physical flag the stage kernels set since the last check (waits for
// the context stream) and applies the context's policy, as PhysicsBase::handle_non_physical_result
// (physics.cpp:298-317): abort_run -> DGX_ERR_NONPHYSICAL (the caller aborts), print_warning ->
// a warning on stderr, return_big_number -> nothing.  *found tells whether the flag was set.
static int check_nonphysical(dgx_ctx *c, const char *where, bool *found = nullptr)
{
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, c->status_dev, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (found) *found = flag != 0;
    if (!flag) return DGX_OK;
    HIPCHK(hipMemsetAsync(c->status_dev, 0, sizeof(int), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    ++c->nonphys_events;
    const std::string msg = std::string("Non-physical result has been detected (negative or NaN density or "
                                        "pressure) in ") + where;
    if (c->nonphys_policy == DGX_ABORT_RUN) return fail(c, DGX_ERR_NONPHYSICAL, msg);
    if (c->nonphys_policy == DGX_PRINT_WARNING) std::fprintf(stderr, "WARNING: %s\n", msg.c_str());
    return DGX_OK;
}

// right-hand side of the stage solution u (aux solve first for Navier-Stokes)
static int residual_into(dgx_ctx *c, const double *u, double *out, int mode)
{
    if (c->P.pde == DGX_NAVIER_STOKES) {
        int rc = launch_aux(c, u);
        if (rc) return rc;
    }
    StageArgs a = base_args(c, u);
    a.mode = mode;
    a.out = out;
    return launch_stage(c, a);
}

int dgx_assemble_residual(dgx_ctx *c)
{
    if (!c) return DGX_ERR_INVALID;
    int rc = residual_into(c, c->u_dev, c->r_dev, MODE_RESIDUAL);
    if (rc) return rc;
    bool found = false;
    rc = check_nonphysical(c, "assemble_residual", &found);
    if (rc == DGX_ERR_NONPHYSICAL) {
        // invalid assembly: right-hand side filled with ones (dg_base.cpp:2850-2875)
        if (launch_fill(c, c->r_dev, c->ndofs, 1.0) != DGX_OK) return DGX_ERR_HIP;
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return rc;
}

int dgx_assemble_residual_device(dgx_ctx *c, const double *u_dev, double *r_dev)
{
    if (!c || !u_dev || !r_dev) return DGX_ERR_INVALID;
    return residual_into(c, u_dev, r_dev, MODE_RESIDUAL);
}

int dgx_rhs_device(dgx_ctx *c, const double *u_dev, double *dudt_dev)
{
    if (!c || !u_dev || !dudt_dev) return DGX_ERR_INVALID;
    return residual_into(c, u_dev, dudt_dev, MODE_RHS);
}

int dgx_apply_inverse_mass_device(dgx_ctx *c, const double *in_dev, double *out_dev)
{
    if (!c || !in_dev || !out_dev) return DGX_ERR_INVALID;
    return launch_inverse_mass(c, in_dev, out_dev);
}

int dgx_apply_inverse_mass_to_rhs(dgx_ctx *c, double *host_out)
{
    // M^-1 of the collocated GLL cell is diagonal, applied on the device
    if (!c || !host_out) return DGX_ERR_INVALID;
    if (!c->minv_dev && hipMalloc(&c->minv_dev, sizeof(double) * c->ndofs) != hipSuccess)
        return fail(c, DGX_ERR_HIP, "hipMalloc failed");
    int rc = launch_inverse_mass(c, c->r_dev, c->minv_dev);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(host_out, c->minv_dev, sizeof(double) * c->ndofs, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DGX_OK;
}

// ---------------------------------------------------------------------------
// explicit RK (Butcher) and low-storage RK tableaus
// runge_kutta_methods.cpp:8-110, low_storage_runge_kutta_methods.cpp:11-220
// ---------------------------------------------------------------------------
struct Butcher { int s; double a[4][4], b[4]; };
struct Lsrk { int s; bool plus; double gamma[11][3], beta[11], delta[11], bhat[11]; };

static bool get_butcher(int m, Butcher &B)
{
    std::memset(&B, 0, sizeof(B));
    switch (m) {
    case DGX_EULER_EX: B.s = 1; B.b[0] = 1.0; return true;
    case DGX_HEUN_EX: B.s = 2; B.a[1][0] = 1.0; B.b[0] = 0.5; B.b[1] = 0.5; return true;
    case DGX_SSPRK3_EX:
        B.s = 3; B.a[1][0] = 1.0; B.a[2][0] = 0.25; B.a[2][1] = 0.25;
        B.b[0] = 1.0 / 6.0; B.b[1] = 1.0 / 6.0; B.b[2] = 2.0 / 3.0; return true;
    case DGX_RK4_EX:
        B.s = 4; B.a[1][0] = 0.5; B.a[2][1] = 0.5; B.a[3][2] = 1.0;
        B.b[0] = 1.0 / 6.0; B.b[1] = 1.0 / 3.0; B.b[2] = 1.0 / 3.0; B.b[3] = 1.0 / 6.0; return true;
    default: return false;
    }
}

static bool get_lsrk(int m, Lsrk &L)
{
    std::memset(&L, 0, sizeof(L));
    if (m == DGX_RK4_3_5_3SSTAR) {
        const double g[6][3] = {{0.0, 0.0, 0.0}, {0.0, 1.0, 0.0},
            {-0.497531095840104, 1.384996869124138, 0.0}, {1.010070514199942, 3.878155713328178, 0.0},
            {-3.196559004608766, -2.324512951813145, 1.642598936063715},
            {1.717835630267259, -0.514633322274467, 0.188295940828347}};
        const double be[6] = {0.0, 0.075152045700771, 0.211361016946069, 1.100713347634329, 0.728537814675568, 0.393172889823198};
        const double de[7] = {1.0, 0.081252332929194, -1.083849060586449, -1.096110881845602, 2.859440022030827,
                              -0.655568367959557, -0.194421504490852};
        L.s = 5; L.plus = false;
        std::memcpy(L.gamma, g, sizeof(g)); std::memcpy(L.beta, be, sizeof(be)); std::memcpy(L.delta, de, sizeof(de));
        return true;
    }
    if (m == DGX_RK3_2_5F_3SSTARPLUS) {
        const double g[6][3] = {{0.0, 0.0, 0.0}, {0.0, 1.0, 0.0},
            {0.2587771979725733308135192812685323706, 0.5528354909301389892439698870483746541, 0.0},
            {-0.1324380360140723382965420909764953437, 0.6731871608203061824849561782794643600, 0.0},
            {0.05056033948190826045833606441415585735, 0.2803103963297672407841316576323901761, 0.2752563273304676380891217287572780582},
            {0.5670532000739313812633197158607642990, 0.5521525447020610386070346724931300367, -0.8950526174674033822276061734289327568}};
        const double be[6] = {0.0, 0.2300298624518076223899418286314123354, 0.3021434166948288809034402119555380003,
                              0.8025606185416310937583009085873554681, 0.4362158943603440930655148245148766471, 0.11292725304550591};
        const double de[5] = {1.0, 0.34076558793345252, 0.34143826550033862, 0.72292753667879872, 0.0};
        const double bh[6] = {0.094841667050357029, 0.17263713394303537, 0.39982431890843712, 0.17180168075801786,
                              0.058819144221557401, 0.1020760551185952388626787099944507877};
        L.s = 5; L.plus = true;
        std::memcpy(L.gamma, g, sizeof(g)); std::memcpy(L.beta, be, sizeof(be));
        std::memcpy(L.delta, de, sizeof(de)); std::memcpy(L.bhat, bh, sizeof(bh));
        return true;
    }
    if (m == DGX_RK4_3_9F_3SSTARPLUS) {
        const double g[10][3] = {{0.0, 0.0, 0.0}, {0.0, 1.0, 0.0},
            {-4.655641447335068552684422206224169103, 2.499262792574495009336242992898153462, 0.0},
            {-0.7720265099645871829248487209517314217, 0.5866820377718875577451517985847920081, 0.0},
            {-4.024436690519806086742256154738379161, 1.205146086523094569925592464380295241, 0.7621006678721315291614677352949377871},
            {-0.02129676284018530966221583708648634733, 0.3474793722186732780030762737753849272, -0.1981182504339400567765766904309673119},
            {-2.435022509790109546199372365866450709, 1.321346060965113109321230804210670518, -0.6228959218699007450469629366684127462},
            {0.01985627297131987000579523283542615256, 0.3119636464694193615946633676950358444, -0.3752248380775956442989480369774937099},
            {-0.2810791146791038566946663374735713961, 0.4351419539684379261368971206040518552, -0.3355438309135169811915662336248989661},
            {0.1689434168754859644351230590422137972, 0.2359698130028753572503744518147537768, -0.04560955005031121479972862973705108039}};
        const double be[10] = {0.0, 0.2836343005184365275160654678626695428, 0.9736500104654741223716056170419660217, 0.3382359225242515288768487569778320563,
                               -0.3584943611106183357043212309791897386, -0.004113944068471528211627210454497620358, 1.427968894048586363415504654313371031,
                               0.01808470948394314017665968411915568633, 0.1605770645946802213926893453819236685, 0.2952227015964591648775833803635147962};
        const double de[9] = {1.0, 1.262923876648114432874834923838556100, 0.7574967189685911558308119415539596711, 0.5163589453140728104667573195005629833,
                              -0.02746327421802609557034437892013640319, -0.4382673178127944142238606608356542890, 1.273587294602656522645691372699677063,
                              -0.6294740283927400326554066998751383342, 0.0};
        const double bh[10] = {0.02483675912451591196775756814283216443, 0.1866327774562103796990092260942180726, 0.05671080795936984495604436622517631183,
                               -0.003447695439149287702616943808570747099, 0.003602245056516636472203469198006404016, 0.4545570622145088936800484247980581766,
                               -0.0002434665289427612407531544765622888855, 0.06642755361103549971517945063138312147, 0.1613697079523505006226025497715177578,
                               0.04955424859358438183052504342394102722};
        L.s = 9; L.plus = true;
        std::memcpy(L.gamma, g, sizeof(g)); std::memcpy(L.beta, be, sizeof(be));
        std::memcpy(L.delta, de, sizeof(de)); std::memcpy(L.bhat, bh, sizeof(bh));
        return true;
    }
    if (m == DGX_RK5_4_10F_3SSTARPLUS) {
        const double g[11][3] = {{0.0, 0.0, 0.0}, {0.0, 1.0, 0.0},
            {0.4043660121685749695640462197806189975, 0.6871467028161416909922221357014564412, 0.0},
            {-0.8503427289575839690883191973980814832, 1.093024748914750833700799552463885117, 0.0},
            {-6.950894175262117526410215315179482885, 3.225975379607193001678365742708874597, -2.393405133244194727221124311276648940},
            {0.9238765192731084931855438934978371889, 1.041153702510101386914019859778740444, -1.902854422421760920850597670305403139},
            {-2.563178056509891340215942413817786020, 1.292821487912164945157744726076279306, -2.820042207399977261483046412236557428},
            {0.2545744879365226143946122067064118430, 0.7391462755788122847651304143259254381, -1.832698465277380999601896111079977378},
            {0.3125831707411998258746812355492206137, 0.1239129251371800313941948224441873274, -0.2199094483084671192328083958346519535},
            {-0.7007114414440507927791249989236719346, 0.1842753472370123193132193302369345580, -0.4082430635847870963724591602173546218},
            {0.4839621016023833375810172323297465039, 0.05712788998796583446479387686662738843, -0.1377669797880289713535665985132703979}};
        const double be[11] = {0.0, 0.2597883554788674084039539165398464630, 0.01777008889438867858759149597539211023, 0.2481636629715501931294746189266601496,
                               0.7941736871152005775821844297293296135, 0.3885391285642019129575902994397298066, 0.1455051657916305055730603387469193768,
                               0.1587517385964749337690916959584348979, 0.1650605617880053419242434594242509601, 0.2118093284937153836908655490906875007,
                               0.1559392342362059886106995325687547506};
        const double de[10] = {1.0, -0.1331778419508803397033287009506932673, 0.8260422814750207498262063505871077303,
                               1.513700425755728332485300719652378197, -1.305810059935023735972298885749903694, 3.036678802924163246003321318996156380,
                               -1.449458274398895177922690618003584514, 3.834313899176362315089976408899373409, 4.122293760012985409330881631526514714,
                               0.0};
        const double bh[11] = {-0.02019255440012066080909442770590267512, 0.02737903480959184339932730854141598275, 0.3028818636145965534365173822296811090,
                               -0.03656843880622222190071445247906780540, 0.3982664774676767729863101188528827405, -0.05715959421140685436681459970502471634,
                               0.09849855103848558320961101178888983150, 0.06654601552456084978615342374581437947, 0.09073479542748112726465375642050504556,
                               0.08432289325330803924891866923939606351, 0.04529095628204896774513180907141004447};
        L.s = 10; L.plus = true;
        std::memcpy(L.gamma, g, sizeof(g)); std::memcpy(L.beta, be, sizeof(be));
        std::memcpy(L.delta, de, sizeof(de)); std::memcpy(L.bhat, bh, sizeof(bh));
        return true;
    }
    return false;
}

int dgx_allocate_ode_system(dgx_ctx *c, int method)
{
    if (!c) return DGX_ERR_INVALID;
    if (c->batch_locked) return fail(c, DGX_ERR_STATE, "instance registers belong to a batch graph");
    Butcher B;
    Lsrk L;
    if (!c->regs.empty()) HIPCHK(hipStreamSynchronize(c->stream));  // the old registers may be in use
    for (double *r : c->regs) dev_free(c, r);
    c->regs.clear();
    int nregs;
    if (get_butcher(method, B)) nregs = 3 + B.s;   // acc, stage x2, k_j (only if referenced)
    else if (get_lsrk(method, L)) nregs = 4;        // S1', S2, S3, S4
    else return fail(c, DGX_ERR_INVALID, "unknown ODE method");
    const size_t bytes = sizeof(double) * (size_t)c->ndofs;
    for (int r = 0; r < nregs; ++r) {
        double *p = static_cast<double *>(dev_alloc(c, bytes));
        if (!p) return fail(c, c->arena ? DGX_ERR_STATE : DGX_ERR_HIP, "RK register allocation failed");
        c->regs.push_back(p);
        HIPCHK(hipMemsetAsync(p, 0, bytes, c->stream));
    }
    c->method = method;
    c->lsrk_cur = 0;
    c->epsilon[0] = c->epsilon[1] = c->epsilon[2] = 1.0;
    c->rrk = false;  // (re-)enabled with dgx_set_relaxation_runge_kutta
    c->rrk_gamma = 1.0;
    return DGX_OK;
}

int dgx_set_track_error_estimate(dgx_ctx *c, int enable)
{
    if (!c) return DGX_ERR_INVALID;
    c->track_error = enable != 0;
    return DGX_OK;
}

// rk_order of the low-storage methods (parameters_ode_solver.cpp:270-296)
static int lsrk_order(int m)
{
    switch (m) {
    case DGX_RK4_3_5_3SSTAR: return 4;
    case DGX_RK3_2_5F_3SSTARPLUS: return 3;
    case DGX_RK4_3_9F_3SSTARPLUS: return 4;
    case DGX_RK5_4_10F_3SSTARPLUS: return 5;
    default: return 0;
    }
}

// Registers of explicit RK: 0 accumulator (solution_update), 1/2 stage
// ping-pong, 3.. stored k_j (only those a later stage reads).
static void butcher_needs_k(const Butcher &B, bool need_k[4])
{
    for (int j = 0; j < 4; ++j) need_k[j] = false;
    for (int j = 0; j < B.s; ++j)
        for (int m = j + 2; m < B.s; ++m)
            if (B.a[m][j] != 0.0) need_k[j] = true;
}

// Stage i of a Butcher step (runge_kutta_ode_solver.cpp:15-99): stage state in,
// k_i = M^-1 R(stage state), accumulator += dt b_i k_i, next stage state
// u_n + dt sum_j a_{i+1,j} k_j.  The last stage writes u_{n+1} over u_n, which it
// no longer reads (its stage state is a stage buffer); a one-stage method reads
// u_n as its stage state, so it keeps the accumulator and the caller swaps.
static int butcher_stage_args(dgx_ctx *c, const Butcher &B, double dt, int i, StageArgs &a)
{
    bool need_k[4];
    butcher_needs_k(B, need_k);
    double *acc = c->regs[0];
    double *stage_buf[2] = {c->regs[1], c->regs[2]};
    const double *cur = (i == 0) ? c->u_dev : stage_buf[(i - 1) & 1];
    a = base_args(c, cur);
    a.mode = MODE_BUTCHER;
    a.dt = dt;
    a.u0 = c->u_dev;
    a.acc = acc;
    a.acc_out = (i == B.s - 1 && B.s > 1) ? c->u_dev : acc;
    a.first_stage = (i == 0);
    a.last_stage = (i == B.s - 1);
    a.stage = i;
    a.b_dt = dt * B.b[i];
    a.k_store = (need_k[i] || c->rrk) ? c->regs[3 + i] : nullptr;
    if (c->rrk) {
        // relaxation RK: the accumulator gathers d = dt sum b_i k_i from zero and u_n is kept
        a.acc_zero = 1;
        a.acc_out = acc;
    }
    if (i < B.s - 1) {
        a.u_next = stage_buf[i & 1];
        for (int j = 0; j <= i; ++j) {
            a.a_next[j] = B.a[i + 1][j];
            a.k_prev[j] = (j < i && need_k[j]) ? c->regs[3 + j] : nullptr;
            if (j < i && B.a[i + 1][j] != 0.0 && !need_k[j])
                return fail(c, DGX_ERR_INVALID, "tableau register bookkeeping");
        }
    }
    return DGX_OK;
}

// Relaxation parameter of a relaxation-RK step (RRKODESolverBase::update_relaxation_parameter,
// rrk_ode_solver_base.cpp:16-35; RootFindingRRKODESolver::compute_relaxation_parameter,
// root_finding_rrk_ode_solver.cpp:16-176): the root of r(g) = S(u_n + g d) - S(u_n) - g est, S the
// integrated numerical entropy, d = dt sum b_i k_i, est = dt sum b_i v(u^(i)) . M k_i; secant
// iteration from the previous parameter +- 1e-5 with the reference's safeguards, bisection if it
// does not converge in 100 iterations.
static int rrk_relaxation(dgx_ctx *c, const Butcher &B, double dt, double *gamma_out)
{
    const int nb = vr_blocks();
    double est = 0.0;
    for (int i = 0; i < B.s; ++i) {
        double ent = 0.0;
        int rc = fetch_partial_sum(c, c->rrk_partial + (size_t)i * nb, &ent);
        if (rc) return rc;
        est += B.b[i] * ent;
    }
    est *= dt;
    const double *un = c->u_dev, *dir = c->regs[0];
    double s_n = 0.0;
    int rc = launch_entropy_combo(c, un, nullptr, 0.0, &s_n);
    if (rc) return rc;
    auto root = [&](double g, double *r) {
        double sg = 0.0;
        int e = launch_entropy_combo(c, un, dir, g, &sg);
        *r = sg - s_n - g * est;
        return e;
    };
    const int limit = 100;
    int iter = 0;
    double gamma_kp1 = 0.0;
    bool failed = false;
    {
        double residual = 1.0, gk = c->rrk_gamma + 1e-5, gkm1 = c->rrk_gamma - 1e-5, rk = 0.0, rkm1 = 0.0;
        if ((rc = root(gk, &rk)) || (rc = root(gkm1, &rkm1))) return rc;
        while (residual > c->rrk_tol && iter < limit) {
            if (rkm1 == rk) {
                gk *= 1.001;
                if ((rc = root(gkm1, &rkm1)) || (rc = root(gk, &rk))) return rc;
            }
            if (gk < 0.5 || gk > 1.5) {
                gk = 1.0;
                if ((rc = root(gk, &rk))) return rc;
            }
            gamma_kp1 = gk - rk * (gk - gkm1) / (rk - rkm1);
            residual = std::fabs(gamma_kp1 - gk);
            iter++;
            gkm1 = gk;
            gk = gamma_kp1;
            rkm1 = rk;
            if ((rc = root(gk, &rk))) return rc;
            if (std::isnan(gk) || std::isnan(gkm1)) {
                gk = 1.0 - 1e-5;
                if ((rc = root(gk, &rk))) return rc;
                gkm1 = 1.0 + 1e-5;
                if ((rc = root(gkm1, &rkm1))) return rc;
                residual = 1.0;
            }
        }
        failed = (iter == limit);
    }
    if (failed) {
        iter = 0;
        double lo = c->rrk_gamma - 0.1, hi = c->rrk_gamma + 0.1, rlo = 0.0, rhi = 0.0, residual = 1.0;
        if ((rc = root(lo, &rlo)) || (rc = root(hi, &rhi))) return rc;
        while (residual > c->rrk_tol && iter < limit) {
            if (rlo * rhi > 0) {
                lo -= 0.1;
                hi += 0.1;
            }
            gamma_kp1 = 0.5 * (lo + hi);
            double r = 0.0;
            if ((rc = root(gamma_kp1, &r))) return rc;
            if (r < 0) {
                lo = gamma_kp1;
                rlo = r;
            } else {
                hi = gamma_kp1;
                rhi = r;
            }
            residual = hi - lo;
            iter++;
        }
    }
    if (iter == limit || !(gamma_kp1 >= 0.5))
        return fail(c, DGX_ERR_NONPHYSICAL, "relaxation RK failed to find a reasonable relaxation parameter");
    *gamma_out = gamma_kp1;
    return DGX_OK;
}

// One Butcher step; *dt_eff is the time advanced (gamma dt with relaxation RK, else dt).
static int step_butcher(dgx_ctx *c, const Butcher &B, double dt, double *dt_eff)
{
    *dt_eff = dt;
    for (int i = 0; i < B.s; ++i) {
        StageArgs a;
        int rc = butcher_stage_args(c, B, dt, i, a);
        if (rc) return rc;
        if (c->P.pde == DGX_NAVIER_STOKES) {
            rc = launch_aux(c, a.u);
            if (rc) return rc;
        }
        rc = launch_stage(c, a);
        if (rc) return rc;
        if (c->rrk) {  // entropy-change estimate of this stage, before its state buffer is reused
            rc = launch_rrk_estimate(c, a.u, a.k_store, c->rrk_partial + (size_t)i * vr_blocks());
            if (rc) return rc;
        }
    }
    if (c->rrk) {
        double gamma = 1.0;
        int rc = rrk_relaxation(c, B, dt, &gamma);
        if (rc) return rc;
        rc = launch_axpy(c, c->u_dev, gamma, c->regs[0], c->u_dev, c->ndofs);  // u_n + gamma d
        if (rc) return rc;
        c->rrk_gamma = gamma;
        *dt_eff = gamma * dt;
        return DGX_OK;
    }
    if (B.s == 1) std::swap(c->u_dev, c->regs[0]); // solution = solution_update
    return DGX_OK;
}

static int step_lsrk(dgx_ctx *c, const Lsrk &L, double dt)
{
    // S1 ping-pong between u_dev and regs[0]; S2 = regs[1]; S3 = regs[2]; S4 = regs[3]
    double *s1[2] = {c->u_dev, c->regs[0]};
    double *S2 = c->regs[1], *S3 = c->regs[2], *S4 = c->regs[3];
    const bool track = L.plus && c->track_error;
    int cur = 0;
    for (int i = 0; i < L.s; ++i) {
        if (c->P.pde == DGX_NAVIER_STOKES) {
            int rc = launch_aux(c, s1[cur]);
            if (rc) return rc;
        }
        StageArgs a = base_args(c, s1[cur]);
        a.mode = MODE_LSRK;
        a.s1_out = s1[cur ^ 1];
        a.s2 = S2;
        a.s3 = S3;
        a.s4 = track ? S4 : nullptr;
        a.lsrk_first = (i == 0);
        a.delta = L.delta[i];
        a.g0 = L.gamma[i + 1][0];
        a.g1 = L.gamma[i + 1][1];
        a.g2 = L.gamma[i + 1][2];
        a.beta = L.beta[i + 1];
        a.bhat = L.bhat[i];
        a.dt = dt;
        int rc = launch_stage(c, a);
        if (rc) return rc;
        cur ^= 1;
    }
    if (track) {
        // FSAL residual for the embedded estimate (low_storage_runge_kutta_ode_solver.cpp:63-69):
        // S4 += bhat_s * dt * M^-1 R(S1); it does not change the solution.
        if (c->P.pde == DGX_NAVIER_STOKES) {
            int rc = launch_aux(c, s1[cur]);
            if (rc) return rc;
        }
        StageArgs a = base_args(c, s1[cur]);
        a.mode = MODE_ACCUM;
        a.out = S4;
        a.bhat = L.bhat[L.s];
        a.dt = dt;
        int rc = launch_stage(c, a);
        if (rc) return rc;
    }
    if (cur == 1) std::swap(c->u_dev, c->regs[0]);
    return DGX_OK;
}

int dgx_step_in_time(dgx_ctx *c, double dt, int n_steps)
{
    if (!c) return DGX_ERR_INVALID;
    if (c->method < 0) return fail(c, DGX_ERR_STATE, "allocate_ode_system first");
    Butcher B;
    Lsrk L;
    const bool isb = get_butcher(c->method, B);
    if (!isb) get_lsrk(c->method, L);
    for (int s = 0; s < n_steps; ++s) {
        double dt_eff = dt;
        int rc = isb ? step_butcher(c, B, dt, &dt_eff) : step_lsrk(c, L, dt);
        if (rc) return rc;
        ++c->iteration;
        c->current_time += dt_eff;
    }
    return check_nonphysical(c, "step_in_time");
}

int dgx_set_relaxation_runge_kutta(dgx_ctx *c, int enable, double root_tolerance)
{
    if (!c || !(root_tolerance > 0.0)) return DGX_ERR_INVALID;
    Butcher B;
    if (enable && (c->method < 0 || !get_butcher(c->method, B)))
        return fail(c, DGX_ERR_STATE, "relaxation RK needs an explicit Butcher method (allocate_ode_system first)");
    if (enable && c->P.pde != DGX_EULER && c->P.pde != DGX_NAVIER_STOKES)
        return fail(c, DGX_ERR_UNSUPPORTED, "relaxation RK is for Euler / Navier-Stokes");
    if (enable && !c->rrk_partial) {
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, sizeof(double) * 4 * (size_t)vr_blocks()));
        c->rrk_partial = static_cast<double *>(p);
    }
    c->rrk = enable != 0;
    c->rrk_tol = root_tolerance;
    c->rrk_gamma = 1.0;
    return DGX_OK;
}

double dgx_relaxation_parameter(const dgx_ctx *c) { return c ? c->rrk_gamma : 0.0; }

int dgx_set_error_adaptive_parameters(dgx_ctx *c, double atol, double rtol, double beta1, double beta2, double beta3)
{
    if (!c || !(atol >= 0.0) || !(rtol >= 0.0) || !(atol + rtol > 0.0)) return DGX_ERR_INVALID;
    c->atol = atol;
    c->rtol = rtol;
    c->beta_pid[0] = beta1;
    c->beta_pid[1] = beta2;
    c->beta_pid[2] = beta3;
    c->epsilon[0] = c->epsilon[1] = c->epsilon[2] = 1.0;  // as constructed (low_storage_..._solver.cpp:12)
    c->track_error = true;  // 3S*+ methods: the FSAL residual feeds the embedded register
    return DGX_OK;
}

int dgx_get_automatic_initial_step_size(dgx_ctx *c, double *dt)
{
    if (!c || !dt) return DGX_ERR_INVALID;
    Lsrk L;
    if (c->method < 0 || !get_lsrk(c->method, L)) return fail(c, DGX_ERR_STATE, "needs a low-storage RK system");
    // the registers S2..S4 are rewritten by the next step before they are read: scratch here
    double *un = c->regs[1], *f0 = c->regs[2], *f1 = c->regs[3];
    double d0 = 0.0, d1 = 0.0, d2 = 0.0, h0, h1;
    int rc = launch_maxabs(c, c->u_dev, nullptr, c->ndofs, &d0);
    if (!rc) rc = residual_into(c, c->u_dev, f0, MODE_RHS);
    if (!rc) rc = launch_maxabs(c, f0, nullptr, c->ndofs, &d1);
    if (rc) return rc;
    h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    rc = launch_axpy(c, c->u_dev, h0, f0, un, c->ndofs);
    if (!rc) rc = residual_into(c, un, f1, MODE_RHS);
    if (!rc) rc = launch_maxabs(c, f1, f0, c->ndofs, &d2);
    if (rc) return rc;
    d2 /= h0;
    if (std::max(d1, d2) <= 1e-15) h1 = std::max(1e-6, h0 * 1e-3);
    else h1 = 0.01 / std::max(d1, d2);
    *dt = std::min(100 * h0, h1);
    return DGX_OK;
}

int dgx_get_error_adaptive_step_size(dgx_ctx *c, double dt, double *dt_next)
{
    if (!c || !dt_next) return DGX_ERR_INVALID;
    Lsrk L;
    if (c->method < 0 || !get_lsrk(c->method, L)) return fail(c, DGX_ERR_STATE, "needs a low-storage RK system");
    if (L.plus && !c->track_error)
        return fail(c, DGX_ERR_STATE, "3S*+ error estimate needs dgx_set_error_adaptive_parameters before the step");
    const int num_delta = (c->method == DGX_RK4_3_5_3SSTAR) ? 7 : L.s;  // parameters_ode_solver.cpp:270-296
    double sum_delta = 0.0;
    for (int i = 0; i < num_delta; ++i) sum_delta += L.delta[i];
    double sum = 0.0;
    // after the step the solution S1 is in u_dev; S2 = regs[1], S3 = regs[2], S4 = regs[3]
    int rc = launch_lsrk_error(c, c->u_dev, L.plus ? c->regs[3] : c->regs[1], c->regs[2], L.plus ? 1 : 0,
                               L.delta[L.s], L.delta[L.s + 1], sum_delta, c->atol, c->rtol, &sum);
    if (rc) return rc;
    const double w = std::pow(sum / (double)c->ndofs, 0.5);
    const double k = lsrk_order(c->method);
    c->epsilon[2] = c->epsilon[1];
    c->epsilon[1] = c->epsilon[0];
    c->epsilon[0] = 1.0 / w;
    *dt_next = std::pow(c->epsilon[0], 1.0 * c->beta_pid[0] / k) * std::pow(c->epsilon[1], 1.0 * c->beta_pid[1] / k) *
               std::pow(c->epsilon[2], 1.0 * c->beta_pid[2] / k) * dt;
    if (!std::isfinite(*dt_next) || !(*dt_next > 0.0)) {
        ++c->nonphys_events;
        return fail(c, DGX_ERR_NONPHYSICAL, "error-adaptive time step is not a positive finite number");
    }
    return DGX_OK;
}

double dgx_current_time(const dgx_ctx *c) { return c ? c->current_time : 0.0; }
int64_t dgx_current_iteration(const dgx_ctx *c) { return c ? c->iteration : 0; }

double dgx_constant_time_step(const dgx_ctx *c, double cfl)
{
    const unsigned ndofs_per_state = (unsigned)(c->ndofs / c->ns);
    const double dx = (c->P.domain_right - c->P.domain_left) / std::pow((double)ndofs_per_state, (1.0 / c->dim));
    return cfl * dx;
}

int dgx_set_time(dgx_ctx *c, double time, int64_t iteration)
{
    if (!c || iteration < 0) return DGX_ERR_INVALID;
    c->current_time = time;
    c->iteration = iteration;
    return DGX_OK;
}

// Initial conditions interpolated at the GLL support points of each cell
// (set_initial_condition.cpp:45-51): TGV (initial_condition_function.cpp:261-336),
// isentropic vortex with parameterised strength/advection (:643-690), evaluated on the device.
int dgx_set_initial_condition_tgv(dgx_ctx *c, int isothermal)
{
    if (!c) return DGX_ERR_INVALID;
    if (c->dim != 3) return fail(c, DGX_ERR_INVALID, "Taylor-Green vortex is 3D");
    int rc = launch_ic(c, 0, isothermal, 0.0, 0.0, 0.0);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->current_time = 0.0;
    c->iteration = 0;
    return DGX_OK;
}

int dgx_set_initial_condition_vortex(dgx_ctx *c, double strength, double u_inf, double v_inf)
{
    if (!c) return DGX_ERR_INVALID;
    if (c->dim != 2) return fail(c, DGX_ERR_INVALID, "isentropic vortex is 2D");
    int rc = launch_ic(c, 1, 0, strength, u_inf, v_inf);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->current_time = 0.0;
    c->iteration = 0;
    return DGX_OK;
}

int dgx_integrated_quantities(dgx_ctx *c, int overint, double *out4)
{
    if (!c || !out4) return DGX_ERR_INVALID;
    return launch_integrated(c, c->u_dev, overint, out4);
}

int dgx_integrated_quantities_static(dgx_ctx *c, const double *u_dev, int overint, double *out4_dev)
{
    if (!c || !out4_dev) return DGX_ERR_INVALID;
    return launch_integrated_static(c, u_dev ? u_dev : c->u_dev, overint, out4_dev);
}

int dgx_numerical_entropy(dgx_ctx *c, double *out)
{
    if (!c || !out) return DGX_ERR_INVALID;
    return launch_entropy(c, c->u_dev, out);
}

int dgx_max_wave_speed(dgx_ctx *c, int overint, double *out)
{
    if (!c || !out || overint < 0) return DGX_ERR_INVALID;
    return launch_wave_speed(c, c->u_dev, overint, out);
}

int dgx_adaptive_time_step(dgx_ctx *c, double cfl, double *dt)
{
    if (!c || !dt) return DGX_ERR_INVALID;
    double vmax = 0.0;
    int rc = launch_wave_speed(c, c->u_dev, 10, &vmax);
    if (rc) return rc;
    if (!(vmax > 0.0) || !std::isfinite(vmax)) {
        ++c->nonphys_events;
        return fail(c, DGX_ERR_NONPHYSICAL, "adaptive time step: maximum wave speed is not a positive finite number");
    }
    const unsigned ndofs_per_state = (unsigned)(c->ndofs / c->ns);
    const double dx = (c->P.domain_right - c->P.domain_left) / std::pow((double)ndofs_per_state, (1.0 / c->dim));
    *dt = cfl * dx / vmax;
    return DGX_OK;
}

int dgx_synchronize(dgx_ctx *c)
{
    if (!c) return DGX_ERR_INVALID;
    HIPCHK(hipStreamSynchronize(c->stream));
    return DGX_OK;
}

int dgx_set_non_physical_behavior(dgx_ctx *c, int policy)
{
    if (!c || policy < DGX_RETURN_BIG_NUMBER || policy > DGX_PRINT_WARNING) return DGX_ERR_INVALID;
    c->nonphys_policy = policy;
    return DGX_OK;
}

int dgx_check_state(dgx_ctx *c)
{
    if (!c) return DGX_ERR_INVALID;
    return check_nonphysical(c, "stage evaluation");
}

int64_t dgx_nonphysical_events(const dgx_ctx *c) { return c ? c->nonphys_events : 0; }

int dgx_set_stream(dgx_ctx *c, void *s)
{
    if (!c) return DGX_ERR_INVALID;
    c->stream = (hipStream_t)s;
    return DGX_OK;
}

int dgx_enable_timing(dgx_ctx *c, int enable)
{
    if (!c) return DGX_ERR_INVALID;
    c->timing.enabled = enable != 0;
    if (enable && !c->timing.a) {
        HIPCHK(hipEventCreate(&c->timing.a));
        HIPCHK(hipEventCreate(&c->timing.b));
    }
    c->timing.sum_ms = 0.0;
    c->timing.n = 0;
    return DGX_OK;
}

int dgx_kernel_timing(const dgx_ctx *c, double *avg_ms, int64_t *n)
{
    if (!c) return DGX_ERR_INVALID;
    if (avg_ms) *avg_ms = c->timing.n ? c->timing.sum_ms / c->timing.n : 0.0;
    if (n) *n = c->timing.n;
    return DGX_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Batched independent solves
// ---------------------------------------------------------------------------
namespace {

struct BatchGroup {
    int dim = 0, n = 0, precision = 0;
    int tile = -1;                       // batched tile (batch_tile); -1: per-instance launches
    std::vector<int> inst;               // instance indices
    BatchBlock *blocks = nullptr;        // device block table
    int nblocks = 0;
    PhysConst *pcs = nullptr;            // device, per slot
    std::vector<StageArgs *> args;       // device, per stage: per slot
};

// batched tile chosen per (dim, n, precision) in this process
std::vector<std::pair<long, int>> g_variant_cache;

} // namespace

struct dgx_batch {
    std::vector<dgx_ctx *> inst;
    std::vector<double> dt;
    std::vector<BatchGroup> groups;
    std::vector<int> variant;            // per instance
    int forced_tile = -1;                // dgx_batch_set_tile; -1: autotuned
    int method = -1;
    hipStream_t stream = nullptr;
    std::vector<hipStream_t> side;       // one per group after the first: the groups' launches of a
                                         // stage are independent graph branches
    std::vector<hipEvent_t> events;      // fork/join events of the captured step
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string err;
};

#define BHIPCHK(call)                                                                    \
    do {                                                                                 \
        hipError_t _e = (call);                                                          \
        if (_e != hipSuccess) {                                                          \
            b->err = std::string(#call) + ": " + hipGetErrorString(_e);                   \
            return DGX_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)

static void batch_free_groups(dgx_batch *b)
{
    for (BatchGroup &g : b->groups) {
        (void)hipFree(g.blocks);
        (void)hipFree(g.pcs);
        for (StageArgs *a : g.args) (void)hipFree(a);
    }
    b->groups.clear();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
    b->events.clear();
}

static bool batch_line_eligible(const dgx_ctx *c)
{
    return c->P.use_split_form && c->P.pde == DGX_EULER && !c->has_boundary && !c->curv;
}

// block table of a group for `epb` elements per block: (slot, first element) per element group
static int group_table(dgx_batch *b, const BatchGroup &g, int epb, BatchBlock **out, int *nblocks)
{
    std::vector<BatchBlock> tab;
    for (int slot = 0; slot < (int)g.inst.size(); ++slot) {
        const dgx_ctx *c = b->inst[g.inst[slot]];
        for (int e = 0; e < c->ncells; e += epb) tab.push_back(BatchBlock{slot, e});
    }
    *out = nullptr;
    BHIPCHK(hipMalloc(out, sizeof(BatchBlock) * tab.size()));
    BHIPCHK(hipMemcpy(*out, tab.data(), sizeof(BatchBlock) * tab.size(), hipMemcpyHostToDevice));
    *nblocks = (int)tab.size();
    return DGX_OK;
}

// block table of a group for tile `tile`
static int batch_blocks(dgx_batch *b, BatchGroup &g, int tile)
{
    int epb = 0;
    int rc = launch_line_batch(g.dim, g.n, g.precision, tile, nullptr, 0, nullptr, nullptr, Ops1D{}, nullptr, &epb);
    if (rc) return rc;
    (void)hipFree(g.blocks);
    g.blocks = nullptr;
    rc = group_table(b, g, epb, &g.blocks, &g.nblocks);
    if (rc) return rc;
    g.tile = tile;
    return DGX_OK;
}

// Times every candidate tile (batch_tile) on the group's first RK stage (median of seven
// launches after a warm-up) and keeps the fastest.
static int batch_autotune(dgx_batch *b, BatchGroup &g, const Ops1D &ops)
{
    if (b->forced_tile >= 0 && batch_blocks(b, g, b->forced_tile) == DGX_OK) return DGX_OK;
    const long key = ((long)g.dim * 16 + g.n) * 4 + g.precision;
    for (auto &kv : g_variant_cache)
        if (kv.first == key) return batch_blocks(b, g, kv.second);
    // the group's own first-stage launch: it reads u_n and writes only the accumulator and the
    // next stage state, which the first step recomputes
    const StageArgs *pa = g.args[0];
    hipEvent_t e0, e1;
    BHIPCHK(hipEventCreate(&e0));
    BHIPCHK(hipEventCreate(&e1));
    // candidates timed round-robin (one launch of each per round, the first round a warm-up), so a
    // drift of the clock or of the other groups' caches during tuning biases no candidate
    std::vector<int> cand, table;                 // candidate tiles, index of their block table
    std::vector<std::pair<int, std::pair<BatchBlock *, int>>> tables;  // epb -> (table, blocks)
    for (int tile = 0; tile < batch_tile_count(g.dim); ++tile) {
        int epb = 0;
        if (launch_line_batch(g.dim, g.n, g.precision, tile, nullptr, 0, nullptr, nullptr, Ops1D{}, nullptr, &epb))
            continue;  // tile does not fit this shape
        int t = -1;
        for (size_t k = 0; k < tables.size(); ++k)
            if (tables[k].first == epb) t = (int)k;
        if (t < 0) {
            BatchBlock *tab = nullptr;
            int nb = 0;
            const int rc = group_table(b, g, epb, &tab, &nb);
            if (rc) return rc;
            tables.push_back({epb, {tab, nb}});
            t = (int)tables.size() - 1;
        }
        cand.push_back(tile);
        table.push_back(t);
    }
    constexpr int ROUNDS = 8;
    std::vector<std::vector<float>> ts(cand.size());
    for (int rep = 0; rep < ROUNDS; ++rep) {
        for (size_t k = 0; k < cand.size(); ++k) {
            const auto &tb = tables[table[k]].second;
            BHIPCHK(hipEventRecord(e0, b->stream));
            if (launch_line_batch(g.dim, g.n, g.precision, cand[k], tb.first, tb.second, pa, g.pcs, ops, b->stream,
                                  nullptr))
                return DGX_ERR_HIP;
            BHIPCHK(hipEventRecord(e1, b->stream));
            BHIPCHK(hipEventSynchronize(e1));
            float t = 0.f;
            BHIPCHK(hipEventElapsedTime(&t, e0, e1));
            if (rep) ts[k].push_back(t);
        }
    }
    int best = 0;
    float best_ms = 1e30f;
    for (size_t k = 0; k < cand.size(); ++k) {
        std::sort(ts[k].begin(), ts[k].end());
        const float med = ts[k][ts[k].size() / 2];
        if (med < best_ms) {
            best_ms = med;
            best = cand[k];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (auto &t : tables) (void)hipFree(t.second.first);
    g_variant_cache.push_back({key, best});
    return batch_blocks(b, g, best);
}

extern "C" {

int dgx_batch_create(const dgx_params *params, int n, dgx_batch **out)
{
    if (!params || n < 1 || !out) return DGX_ERR_INVALID;
    *out = nullptr;
    dgx_batch *b = new dgx_batch();
    if (hipStreamCreate(&b->stream) != hipSuccess) {
        delete b;
        return DGX_ERR_HIP;
    }
    for (int i = 0; i < n; ++i) {
        dgx_ctx *c = nullptr;
        int rc = dgx_create(&params[i], &c);
        if (rc) {
            dgx_batch_destroy(b);
            return rc;
        }
        c->stream = b->stream;
        b->inst.push_back(c);
    }
    b->variant.assign(n, 0);
    *out = b;
    return DGX_OK;
}

int dgx_batch_destroy(dgx_batch *b)
{
    if (!b) return DGX_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free_groups(b);
    for (dgx_ctx *c : b->inst) {
        c->stream = nullptr;
        c->batch_locked = false;
        dgx_destroy(c);
    }
    for (hipStream_t st : b->side) (void)hipStreamDestroy(st);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return DGX_OK;
}

const char *dgx_batch_last_error(const dgx_batch *b) { return b ? b->err.c_str() : "null batch"; }
int dgx_batch_size(const dgx_batch *b) { return b ? (int)b->inst.size() : 0; }

dgx_ctx *dgx_batch_instance(dgx_batch *b, int i)
{
    if (!b || i < 0 || i >= (int)b->inst.size()) return nullptr;
    return b->inst[i];
}

int dgx_batch_variant(const dgx_batch *b, int i)
{
    if (!b || i < 0 || i >= (int)b->inst.size()) return -1;
    return b->variant[i];
}

int dgx_batch_set_tile(dgx_batch *b, int tile)
{
    if (!b || tile < -1 || tile >= BATCH_TILES_2D) return DGX_ERR_INVALID;
    b->forced_tile = tile;
    return DGX_OK;
}

int dgx_batch_tile(const dgx_batch *b, int i, int *elements_per_block, int *waves_bound, int *threads_per_block)
{
    if (!b || i < 0 || i >= (int)b->inst.size() || !elements_per_block || !waves_bound || !threads_per_block)
        return DGX_ERR_INVALID;
    *elements_per_block = 0;
    *waves_bound = 0;
    *threads_per_block = 0;
    const int v = b->variant[i];
    if (!v) return DGX_OK;
    const dgx_ctx *c = b->inst[i];
    const bool fast = c->P.precision == DGX_MIXED;
    const BatchTile t = batch_tile(c->dim, fast, v - 1);
    *waves_bound = t.waves;
    return launch_line_batch(c->dim, c->n, c->P.precision, v - 1, nullptr, 0, nullptr, nullptr, Ops1D{}, nullptr,
                             elements_per_block, threads_per_block);
}

int dgx_batch_allocate(dgx_batch *b, int method, const double *dt)
{
    if (!b || !dt) return DGX_ERR_INVALID;
    Butcher B;
    if (!get_butcher(method, B) || B.s < 2) {
        b->err = "batched stepping supports explicit Butcher methods with at least two stages";
        return DGX_ERR_UNSUPPORTED;
    }
    BHIPCHK(hipStreamSynchronize(b->stream));
    batch_free_groups(b);
    const int n = (int)b->inst.size();
    b->dt.assign(dt, dt + n);
    for (int i = 0; i < n; ++i) {
        dgx_ctx *c = b->inst[i];
        c->batch_locked = false;
        int rc = dgx_allocate_ode_system(c, method);
        if (rc) {
            b->err = c->err;
            return rc;
        }
        c->batch_locked = true;
        c->timing.enabled = false;  // no host synchronisation inside the captured graph
    }
    b->method = method;
    // groups of split-form Euler instances sharing (dim, n, precision)
    for (int i = 0; i < n; ++i) {
        dgx_ctx *c = b->inst[i];
        b->variant[i] = 0;
        if (!batch_line_eligible(c)) continue;
        BatchGroup *g = nullptr;
        for (BatchGroup &h : b->groups)
            if (h.dim == c->dim && h.n == c->n && h.precision == c->P.precision) g = &h;
        if (!g) {
            b->groups.emplace_back();
            g = &b->groups.back();
            g->dim = c->dim;
            g->n = c->n;
            g->precision = c->P.precision;
        }
        g->inst.push_back(i);
    }
    for (BatchGroup &g : b->groups) {
        std::vector<PhysConst> pcs;
        for (int i : g.inst) pcs.push_back(b->inst[i]->pc);
        BHIPCHK(hipMalloc(&g.pcs, sizeof(PhysConst) * pcs.size()));
        BHIPCHK(hipMemcpy(g.pcs, pcs.data(), sizeof(PhysConst) * pcs.size(), hipMemcpyHostToDevice));
        for (int st = 0; st < B.s; ++st) {
            std::vector<StageArgs> as(g.inst.size());
            for (size_t slot = 0; slot < g.inst.size(); ++slot) {
                dgx_ctx *c = b->inst[g.inst[slot]];
                int rc = butcher_stage_args(c, B, b->dt[g.inst[slot]], st, as[slot]);
                if (rc) {
                    b->err = c->err;
                    return rc;
                }
            }
            StageArgs *d = nullptr;
            BHIPCHK(hipMalloc(&d, sizeof(StageArgs) * as.size()));
            BHIPCHK(hipMemcpy(d, as.data(), sizeof(StageArgs) * as.size(), hipMemcpyHostToDevice));
            g.args.push_back(d);
        }
        const Ops1D ops = b->inst[g.inst[0]]->ops;
        int rc = batch_autotune(b, g, ops);
        if (rc) return rc;
        for (int i : g.inst) b->variant[i] = g.tile + 1;
    }
    // capture one RK step of every instance; within a stage the groups' launches are branches of
    // the graph (forked from and joined into the batch stream), so the GPU can run blocks of
    // several groups at once
    const int ng = (int)b->groups.size();
    while ((int)b->side.size() < ng - 1) {
        hipStream_t st = nullptr;
        BHIPCHK(hipStreamCreate(&st));
        b->side.push_back(st);
    }
    for (int k = 0; k < B.s * ng; ++k) {
        hipEvent_t e = nullptr;
        BHIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->events.push_back(e);
    }
    BHIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = DGX_OK;
    for (int st = 0; st < B.s && rc == DGX_OK; ++st) {
        hipEvent_t *ev = b->events.data() + st * ng;  // [0]: fork, [k]: join of group k
        if (ng > 1 && hipEventRecord(ev[0], b->stream) != hipSuccess) rc = DGX_ERR_HIP;
        for (int k = 0; k < ng && rc == DGX_OK; ++k) {
            BatchGroup &g = b->groups[k];
            hipStream_t gs = k ? b->side[k - 1] : b->stream;
            if (k && hipStreamWaitEvent(gs, ev[0], 0) != hipSuccess) rc = DGX_ERR_HIP;
            const Ops1D &ops = b->inst[g.inst[0]]->ops;
            if (!rc)
                rc = launch_line_batch(g.dim, g.n, g.precision, g.tile, g.blocks, g.nblocks, g.args[st], g.pcs, ops,
                                       gs, nullptr);
            if (!rc && k && hipEventRecord(ev[k], gs) != hipSuccess) rc = DGX_ERR_HIP;
        }
        for (int k = 1; k < ng && rc == DGX_OK; ++k)
            if (hipStreamWaitEvent(b->stream, ev[k], 0) != hipSuccess) rc = DGX_ERR_HIP;
        if (rc && b->err.empty()) b->err = "capture of the batched step failed";
        for (int i = 0; i < n && rc == DGX_OK; ++i) {
            dgx_ctx *c = b->inst[i];
            if (b->variant[i]) continue;
            StageArgs a;
            rc = butcher_stage_args(c, B, b->dt[i], st, a);
            if (!rc && c->P.pde == DGX_NAVIER_STOKES) rc = launch_aux(c, a.u);
            if (!rc) rc = launch_stage(c, a);
            if (rc) b->err = c->err;
        }
    }
    hipGraph_t graph = nullptr;
    hipError_t ce = hipStreamEndCapture(b->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ce != hipSuccess) {
        b->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(ce);
        return DGX_ERR_HIP;
    }
    b->graph = graph;
    BHIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return DGX_OK;
}

int dgx_batch_step_in_time(dgx_batch *b, int n_steps)
{
    if (!b) return DGX_ERR_INVALID;
    if (!b->exec) {
        b->err = "dgx_batch_allocate first";
        return DGX_ERR_STATE;
    }
    for (int s = 0; s < n_steps; ++s) {
        BHIPCHK(hipGraphLaunch(b->exec, b->stream));
        for (size_t i = 0; i < b->inst.size(); ++i) {
            ++b->inst[i]->iteration;
            b->inst[i]->current_time += b->dt[i];
        }
    }
    return DGX_OK;
}

int dgx_batch_synchronize(dgx_batch *b)
{
    if (!b) return DGX_ERR_INVALID;
    BHIPCHK(hipStreamSynchronize(b->stream));
    int rc = DGX_OK;
    for (dgx_ctx *c : b->inst) {  // every instance's non-physical flag, under its own policy
        const int r = check_nonphysical(c, "batched step");
        if (r && !rc) {
            rc = r;
            b->err = c->err;
        }
    }
    return rc;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Serving: continuous batching of independent requests
// ---------------------------------------------------------------------------
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    dgx_request r{};
    Clock::time_point submitted;
    int64_t bytes = 0;
    dgx_ctx *c = nullptr;   // while active
    double dt = 0.0;        // constant step (last one clipped)
    int64_t steps = 0;
};

struct Result {
    std::vector<double> u;
    double latency = 0.0, dof_updates = 0.0;
    int64_t steps = 0;
    bool nonphysical = false;   // a stage of this request met a non-physical state
};

int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

} // namespace

struct dgx_server {
    int64_t budget = 0, used = 0;
    dgx::Arena arena;               // device memory of the active requests (sized to the budget)
    int max_active = 0, method = -1;
    Butcher B{};
    hipStream_t stream = nullptr;
    int64_t next_id = 1;
    std::vector<Request> queue;     // FIFO
    std::vector<Request> active;
    std::map<int64_t, Result> done;
    char *stage_host = nullptr;     // pinned staging of the per-step tables
    char *stage_dev = nullptr;
    int64_t stage_cap = 0;
    std::string err;
};

static int64_t request_ndofs(const dgx_params &p)
{
    const int64_t n = p.poly_degree + 1, N = p.n_cells_1d;
    const int64_t np = (p.dim == 3) ? n * n * n : n * n, ncells = (p.dim == 3) ? N * N * N : N * N;
    return ncells * np * (p.dim + 2);
}

#define SHIPCHK(call)                                                                    \
    do {                                                                                 \
        hipError_t _e = (call);                                                          \
        if (_e != hipSuccess) {                                                          \
            s->err = std::string(#call) + ": " + hipGetErrorString(_e);                   \
            return DGX_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)

// Admits queued requests first come first served while a slot is free and their footprint fits
// the budget.  Their buffers come from the server's arena and their connectivity and initial
// data are computed by kernels on the server stream, so admission neither allocates device
// memory nor waits for the requests already on the device.
static int server_admit(dgx_server *s)
{
    while (!s->queue.empty() && (int)s->active.size() < s->max_active && s->used + s->queue.front().bytes <= s->budget) {
        Request q = s->queue.front();
        dgx_ctx *c = nullptr;
        int rc = create_ctx(&q.r.params, &s->arena, s->stream, &c);
        if (rc == DGX_ERR_STATE) break;  // no contiguous range free in the arena yet: wait for retirements
        if (rc) {
            s->err = "context set-up failed for an admitted request";
            return rc;
        }
        rc = (q.r.initial_condition == DGX_IC_VORTEX)
                 ? launch_ic(c, 1, 0, q.r.vortex_strength, q.r.vortex_u, q.r.vortex_v)
                 : ((c->dim == 3) ? launch_ic(c, 0, q.r.tgv_isothermal, 0.0, 0.0, 0.0)
                                  : fail(c, DGX_ERR_INVALID, "Taylor-Green vortex is 3D"));
        if (!rc) rc = dgx_allocate_ode_system(c, s->method);
        if (rc == DGX_ERR_STATE && c->arena) {  // registers did not fit: undo, wait for retirements
            dgx_destroy(c);
            break;
        }
        if (rc) {
            s->err = c->err;
            dgx_destroy(c);
            return rc;
        }
        s->queue.erase(s->queue.begin());
        const double L = q.r.params.domain_right - q.r.params.domain_left;
        const double dx = L / std::pow((double)(unsigned)(c->ndofs / c->ns), 1.0 / c->dim);
        q.dt = q.r.cfl * dx / q.r.wave_speed;
        q.c = c;
        s->used += q.bytes;
        s->active.push_back(q);
    }
    return DGX_OK;
}

// One RK step of every active request; stage arguments of the split-form Euler
// requests go to the device in one packed copy, then one batched launch per
// (dim, p, precision) group per stage; other schemes launch per request.
static int server_advance(dgx_server *s)
{
    const int na = (int)s->active.size();
    const int S = s->B.s;
    std::vector<double> step(na);
    std::vector<std::vector<StageArgs>> args(na, std::vector<StageArgs>(S));
    for (int k = 0; k < na; ++k) {
        Request &q = s->active[k];
        double h = q.dt;
        if (q.c->current_time + h > q.r.final_time) h = q.r.final_time - q.c->current_time;
        step[k] = h;
        for (int st = 0; st < S; ++st) {
            int rc = butcher_stage_args(q.c, s->B, h, st, args[k][st]);
            if (rc) {
                s->err = q.c->err;
                return rc;
            }
        }
    }
    struct G {
        int dim, n, precision, tile, epb;
        std::vector<int> members;
        int64_t off_blocks, off_pcs, off_args;
        int nblocks;
    };
    std::vector<G> groups;
    std::vector<bool> grouped(na, false);
    for (int k = 0; k < na; ++k) {
        dgx_ctx *c = s->active[k].c;
        if (!batch_line_eligible(c)) continue;
        G *g = nullptr;
        for (G &x : groups)
            if (x.dim == c->dim && x.n == c->n && x.precision == c->P.precision) g = &x;
        if (!g) {
            G x{};
            x.dim = c->dim;
            x.n = c->n;
            x.precision = c->P.precision;
            const long key = ((long)x.dim * 16 + x.n) * 4 + x.precision;
            x.tile = 0;
            for (auto &kv : g_variant_cache)
                if (kv.first == key) x.tile = kv.second;
            if (launch_line_batch(x.dim, x.n, x.precision, x.tile, nullptr, 0, nullptr, nullptr, Ops1D{}, nullptr,
                                  &x.epb) != DGX_OK) {
                x.tile = 0;
                launch_line_batch(x.dim, x.n, x.precision, 0, nullptr, 0, nullptr, nullptr, Ops1D{}, nullptr, &x.epb);
            }
            groups.push_back(x);
            g = &groups.back();
        }
        g->members.push_back(k);
        grouped[k] = true;
    }
    // pack the tables
    int64_t bytes = 0;
    for (G &g : groups) {
        g.nblocks = 0;
        for (int k : g.members) g.nblocks += (s->active[k].c->ncells + g.epb - 1) / g.epb;
        g.off_blocks = bytes;
        bytes = align256(bytes + sizeof(BatchBlock) * g.nblocks);
        g.off_pcs = bytes;
        bytes = align256(bytes + sizeof(PhysConst) * g.members.size());
        g.off_args = bytes;
        bytes = align256(bytes + sizeof(StageArgs) * g.members.size() * S);
    }
    if (bytes > s->stage_cap) {
        (void)hipHostFree(s->stage_host);
        (void)hipFree(s->stage_dev);
        s->stage_host = nullptr;
        s->stage_dev = nullptr;
        const int64_t cap = std::max<int64_t>(bytes * 2, 1 << 20);
        SHIPCHK(hipHostMalloc((void **)&s->stage_host, cap, hipHostMallocDefault));
        SHIPCHK(hipMalloc((void **)&s->stage_dev, cap));
        s->stage_cap = cap;
    }
    for (G &g : groups) {
        BatchBlock *bb = (BatchBlock *)(s->stage_host + g.off_blocks);
        PhysConst *pp = (PhysConst *)(s->stage_host + g.off_pcs);
        StageArgs *aa = (StageArgs *)(s->stage_host + g.off_args);
        int nb = 0;
        const int m = (int)g.members.size();
        for (int slot = 0; slot < m; ++slot) {
            dgx_ctx *c = s->active[g.members[slot]].c;
            for (int e = 0; e < c->ncells; e += g.epb) bb[nb++] = BatchBlock{slot, e};
            pp[slot] = c->pc;
            for (int st = 0; st < S; ++st) aa[st * m + slot] = args[g.members[slot]][st];
        }
    }
    if (bytes) SHIPCHK(hipMemcpyAsync(s->stage_dev, s->stage_host, bytes, hipMemcpyHostToDevice, s->stream));
    for (int st = 0; st < S; ++st) {
        for (G &g : groups) {
            const int m = (int)g.members.size();
            const Ops1D &ops = s->active[g.members[0]].c->ops;
            int rc = launch_line_batch(g.dim, g.n, g.precision, g.tile, (const BatchBlock *)(s->stage_dev + g.off_blocks),
                                       g.nblocks, (const StageArgs *)(s->stage_dev + g.off_args) + st * m,
                                       (const PhysConst *)(s->stage_dev + g.off_pcs), ops, s->stream, nullptr);
            if (rc) {
                s->err = "batched stage launch failed";
                return rc;
            }
        }
        for (int k = 0; k < na; ++k) {
            if (grouped[k]) continue;
            dgx_ctx *c = s->active[k].c;
            int rc = DGX_OK;
            if (c->P.pde == DGX_NAVIER_STOKES) rc = launch_aux(c, args[k][st].u);
            if (!rc) rc = launch_stage(c, args[k][st]);
            if (rc) {
                s->err = c->err;
                return rc;
            }
        }
    }
    SHIPCHK(hipStreamSynchronize(s->stream));  // also releases the staging buffer
    for (int k = 0; k < na; ++k) {
        Request &q = s->active[k];
        ++q.c->iteration;
        q.c->current_time += step[k];
        ++q.steps;
    }
    return DGX_OK;
}

static int server_retire(dgx_server *s)
{
    for (size_t k = 0; k < s->active.size();) {
        Request &q = s->active[k];
        if (q.c->current_time < q.r.final_time) {
            ++k;
            continue;
        }
        Result res;
        res.u.resize((size_t)q.c->ndofs);
        SHIPCHK(hipMemcpyAsync(res.u.data(), q.c->u_dev, sizeof(double) * q.c->ndofs, hipMemcpyDeviceToHost,
                               s->stream));
        SHIPCHK(hipStreamSynchronize(s->stream));
        bool found = false;
        if (check_nonphysical(q.c, "served request", &found) == DGX_ERR_HIP) {
            s->err = q.c->err;
            return DGX_ERR_HIP;
        }
        res.nonphysical = found;
        res.latency = std::chrono::duration<double>(Clock::now() - q.submitted).count();
        res.steps = q.steps;
        res.dof_updates = (double)q.c->ndofs * s->B.s * q.steps;
        s->done[q.id] = std::move(res);
        dgx_destroy(q.c);  // buffers back to the arena
        s->used -= q.bytes;
        s->active.erase(s->active.begin() + k);
    }
    return DGX_OK;
}

extern "C" {

int64_t dgx_request_n_dofs(const dgx_request *r) { return r ? request_ndofs(r->params) : 0; }

int64_t dgx_request_bytes(const dgx_request *r, int method)
{
    if (!r) return 0;
    Butcher B;
    const int nregs = get_butcher(method, B) ? 3 + B.s : 4;
    const int64_t nd = request_ndofs(r->params);
    const int64_t N = r->params.n_cells_1d, ncells = (r->params.dim == 3) ? N * N * N : N * N;
    int64_t b = (int64_t)sizeof(double) * nd * (2 + nregs) + (int64_t)sizeof(int) * ncells * 2 * r->params.dim;
    if (r->params.pde == DGX_NAVIER_STOKES) b += (int64_t)sizeof(double) * nd * r->params.dim;
    return b;
}

int dgx_server_create(int64_t hbm_budget_bytes, int max_active, int method, dgx_server **out)
{
    if (!out || hbm_budget_bytes <= 0 || max_active < 1) return DGX_ERR_INVALID;
    *out = nullptr;
    Butcher B;
    if (!get_butcher(method, B) || B.s < 2) return DGX_ERR_UNSUPPORTED;
    dgx_server *s = new dgx_server();
    s->budget = hbm_budget_bytes;
    s->max_active = max_active;
    s->method = method;
    s->B = B;
    if (hipStreamCreate(&s->stream) != hipSuccess) {
        delete s;
        return DGX_ERR_HIP;
    }
    // the budget plus room for the per-request status words and 256-byte alignment of each buffer
    if (s->arena.init(hbm_budget_bytes + (16ll << 20) + 8192ll * max_active) != DGX_OK) {
        (void)hipStreamDestroy(s->stream);
        delete s;
        return DGX_ERR_HIP;
    }
    *out = s;
    return DGX_OK;
}

int dgx_server_destroy(dgx_server *s)
{
    if (!s) return DGX_OK;
    (void)hipStreamSynchronize(s->stream);
    for (Request &q : s->active) dgx_destroy(q.c);
    s->arena.release();
    (void)hipHostFree(s->stage_host);
    (void)hipFree(s->stage_dev);
    (void)hipStreamDestroy(s->stream);
    delete s;
    return DGX_OK;
}

const char *dgx_server_last_error(const dgx_server *s) { return s ? s->err.c_str() : "null server"; }

int dgx_server_submit(dgx_server *s, const dgx_request *r, int64_t *id)
{
    if (!s || !r || !id) return DGX_ERR_INVALID;
    Request q;
    q.r = *r;
    q.bytes = dgx_request_bytes(r, s->method);
    if (q.bytes > s->budget || !(r->final_time > 0.0) || !(r->cfl > 0.0) || !(r->wave_speed > 0.0)) {
        s->err = "request cannot be served (exceeds the HBM budget or has no positive time step / final time)";
        return DGX_ERR_INVALID;
    }
    q.id = s->next_id++;
    q.submitted = Clock::now();
    s->queue.push_back(q);
    *id = q.id;
    return DGX_OK;
}

int dgx_server_step(dgx_server *s, int *n_active, int *n_queued, int *n_done)
{
    if (!s) return DGX_ERR_INVALID;
    int rc = server_admit(s);
    if (!rc && !s->active.empty()) rc = server_advance(s);
    if (!rc) rc = server_retire(s);
    if (n_active) *n_active = (int)s->active.size();
    if (n_queued) *n_queued = (int)s->queue.size();
    if (n_done) *n_done = (int)s->done.size();
    return rc;
}

int dgx_server_result(dgx_server *s, int64_t id, double *host_u, double *latency_s, int64_t *steps,
                      double *dof_updates)
{
    if (!s) return DGX_ERR_INVALID;
    auto it = s->done.find(id);
    if (it == s->done.end()) return DGX_ERR_STATE;
    if (host_u) std::memcpy(host_u, it->second.u.data(), sizeof(double) * it->second.u.size());
    if (latency_s) *latency_s = it->second.latency;
    if (steps) *steps = it->second.steps;
    if (dof_updates) *dof_updates = it->second.dof_updates;
    if (it->second.nonphysical) {
        s->err = "request met a non-physical state (negative or NaN density or pressure)";
        return DGX_ERR_NONPHYSICAL;
    }
    return DGX_OK;
}

int dgx_server_release(dgx_server *s, int64_t id)
{
    if (!s) return DGX_ERR_INVALID;
    return s->done.erase(id) ? DGX_OK : DGX_ERR_STATE;
}

int dgx_server_memory(const dgx_server *s, int64_t *in_use, int64_t *budget)
{
    if (!s) return DGX_ERR_INVALID;
    if (in_use) *in_use = s->used;
    if (budget) *budget = s->budget;
    return DGX_OK;
}

} // extern "C"
