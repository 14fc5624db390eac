/*
 * dg_oracle.c -- TEST INFRASTRUCTURE.  CPU FP64 restatement of PHiLiP's
 * explicit strong-form DG/FR residual and explicit RK time step.  It is the
 * parity checker for the HIP product path and the CPU baseline leg of the
 * benchmark; no product code links or calls it.
 *
 * Every function names the reference file:line it restates.  The loop
 * structure follows the reference: per cell a volume term, then the faces the
 * cell owns (lower index does the work), the neighbour's share of a face is
 * scattered straight into the global right-hand side, and the inverse mass is
 * a separate pass (src/dg/dg_base.cpp:453-870, :2820-2843, :2282-2284,
 * :4089-4260).  Metric terms are those of an affine Cartesian cell of side h:
 * J = h I, det J = h^d, cofactor C = h^(d-1) I (the values the reference's
 * metric_operators produce for straight_periodic_cube up to round-off,
 * src/operators/operators.cpp:2340-2884).
 *
 * Parity of this restatement is pinned by the reference's own golden values
 * (viscous TGV kinetic energy and dissipation rate at 1e-10, see
 * tests/test_oracle_golden.py) and by closed-form / invariant checks
 * (tests/test_oracle_*.py).  Parity unpinned: penalty_scaling's cell extent on
 * curved cells (the longest cell edge along the direction, deal.II's
 * extent_in_direction) -- pinned only on parallelepipeds, where it is h.
 *
 * Curvilinear grids (node coordinates set by dgo_set_grid_nodes or the curved
 * periodic grid) replace the affine metric by per-node metric terms.
 */
#include "dg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN DGO_MAXN
#define NS_MAX 5

struct dgo_ctx {
    dgo_params P;
    int dim, p, n, nq, Np, nfp, ns, ncells;
    double h, detJ, cof;
    /* solution support points (GLL, deal.II FE_DGQ) and 1D quadrature */
    double xs[MAXN], ws[MAXN];
    double xq[MAXN], wq[MAXN];
    /* operators, row-major */
    double chi[MAXN * MAXN];      /* [nq][n]  soln basis at quad nodes   */
    double chi_grad[MAXN * MAXN]; /* [nq][n]                              */
    double chi_surf[2][MAXN];     /* [1][n]   soln basis at x=0 / x=1     */
    double fb_vol[MAXN * MAXN];   /* [nq][nq] flux basis (Lagrange on quad nodes) */
    double fb_grad[MAXN * MAXN];
    double fb_surf[2][MAXN];
    double stiff[MAXN * MAXN];    /* flux-basis stiffness  S_ij = sum_q phi_i phi_j' w_q */
    double skew[MAXN * MAXN];     /* S - S^T */
    double mass[MAXN * MAXN];     /* soln mass on quad nodes */
    double minv[MAXN * MAXN];
    double proj[MAXN * MAXN];     /* [n][nq]  M^-1 chi^T W */
    double one1[1];
    double *wvol;                 /* [Np] volume quad weights */
    double *wface;                /* [nfp] face quad weights */
    double *ones_vol, *ones_face;
    unsigned *had_rows, *had_cols;            /* volume Hadamard pattern [Np*n][dim] */
    unsigned *had_srows[3], *had_scols[3];    /* surface patterns per normal dir */
    int *cell_ijk;                /* [ncells][3] */
    int *nbr;                     /* [ncells][2*dim] */
    double *aux[3];               /* lifted gradient (NS), DoF order */
    double *aux_rhs[3];
    /* physics constants (src/physics/euler.cpp:13-67, navier_stokes.cpp:16-55) */
    double gam, gamm1, mach2, pinf, temperature_ratio;
    double vinf[3];               /* free-stream velocity (Euler constructor, euler.cpp:53-61) */
    /* geometry: node coordinates [cell][node][3]; on a curvilinear grid (curv) the metric cofactor
       [cell][node][3][3] and det J [cell][node] at the GLL nodes (metric_operators,
       operators.cpp:2520-2884) */
    double *xn, *metC, *metJ;
    int curv;
    int fsal;                     /* 3S*+ methods: evaluate the FSAL residual (reference: always) */
};

static int g_threads = 1;
void dgo_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

/* Planted mistakes for the tests that check the oracle's own checks (DGO_MUT_*): a test
   switches one on, shows that the identity it relies on then fails, and switches it off. */
static int g_mutation = DGO_MUT_NONE;
void dgo_set_mutation(int m) { g_mutation = m; }

/* ------------------------------------------------------------------------ */
/* 1D quadrature (deal.II QGauss / QGaussLobatto on [0,1])                  */
/* ------------------------------------------------------------------------ */
static void legendre(int n, double x, double *pn, double *dpn)
{
    /* P_n and P_n' on [-1,1] by the three-term recurrence */
    double p0 = 1.0, p1 = x;
    if (n == 0) { *pn = 1.0; *dpn = 0.0; return; }
    for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1; p1 = p2;
    }
    *pn = p1;
    *dpn = (fabs(x) < 1.0) ? n * (x * p1 - p0) / (x * x - 1.0) : 0.5 * n * (n + 1.0) * pow(x, n + 1);
}

void dgo_gauss_legendre(int n, double *x, double *w)
{
    /* Newton on P_n from Chebyshev guesses; mapped to [0,1] (weights sum 1); n <= 64 (the
       over-integration rules QGauss(p + 1 + overint) exceed the basis size) */
    double xt[64], wt[64];
    if (n > 64) { fprintf(stderr, "dgo_gauss_legendre: at most 64 points\n"); abort(); }
    for (int i = 0; i < n; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double pn, dpn;
            legendre(n, z, &pn, &dpn);
            double dz = pn / dpn;
            z -= dz;
            if (fabs(dz) < 1e-17) break;
        }
        double pn, dpn;
        legendre(n, z, &pn, &dpn);
        xt[i] = z;
        wt[i] = 2.0 / ((1.0 - z * z) * dpn * dpn);
    }
    for (int i = 0; i < n; ++i) { /* ascending */
        x[i] = 0.5 * (1.0 - xt[i]);
        w[i] = 0.5 * wt[i];
    }
    for (int i = 0; i < n / 2; ++i) { /* symmetrize */
        double a = 0.5 * (x[i] + 1.0 - x[n - 1 - i]);
        x[i] = a; x[n - 1 - i] = 1.0 - a;
        double b = 0.5 * (w[i] + w[n - 1 - i]);
        w[i] = b; w[n - 1 - i] = b;
    }
    if (n % 2 == 1) x[n / 2] = 0.5;
}

void dgo_gauss_lobatto(int n, double *x, double *w)
{
    /* end points and the roots of P'_N (N = n-1) by Newton with P''_N from the Legendre ODE;
       weights 2/(N(N+1) P_N^2) ; mapped to [0,1] */
    const int N = n - 1;
    double xt[MAXN];
    for (int i = 0; i < n; ++i) {
        double z = cos(M_PI * i / N);
        if (i > 0 && i < N) {
            for (int it = 0; it < 100; ++it) {
                double pn, dpn;
                legendre(N, z, &pn, &dpn);
                const double d2 = (2.0 * z * dpn - N * (N + 1.0) * pn) / (1.0 - z * z);
                const double dz = dpn / d2;
                z -= dz;
                if (fabs(dz) < 1e-17) break;
            }
        }
        xt[i] = z;
    }
    xt[0] = 1.0; xt[N] = -1.0;
    for (int i = 0; i < n; ++i) {
        double pn, dpn;
        legendre(N, xt[i], &pn, &dpn);
        x[i] = 0.5 * (1.0 - xt[i]);
        w[i] = 0.5 * 2.0 / (N * (N + 1.0) * pn * pn);
    }
    for (int i = 0; i < n / 2; ++i) {
        double a = 0.5 * (x[i] + 1.0 - x[n - 1 - i]);
        x[i] = a; x[n - 1 - i] = 1.0 - a;
        double b = 0.5 * (w[i] + w[n - 1 - i]);
        w[i] = b; w[n - 1 - i] = b;
    }
    x[0] = 0.0; x[n - 1] = 1.0;
    if (n % 2 == 1) x[n / 2] = 0.5;
}

void dgo_lagrange_values(int n, const double *s, double x, double *val, double *der)
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

static void invert(int n, const double *A, double *Ainv)
{
    double a[MAXN][2 * MAXN];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 2 * n; ++j)
            a[i][j] = (j < n) ? A[i * n + j] : (j - n == i ? 1.0 : 0.0);
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (piv != c)
            for (int j = 0; j < 2 * n; ++j) { double t = a[c][j]; a[c][j] = a[piv][j]; a[piv][j] = t; }
        double d = a[c][c];
        for (int j = 0; j < 2 * n; ++j) a[c][j] /= d;
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = a[r][c];
            if (f != 0.0)
                for (int j = 0; j < 2 * n; ++j) a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Ainv[i * n + j] = a[i][n + j];
}

/* ------------------------------------------------------------------------ */
/* sum factorisation  (src/operators/operators.cpp:294-398, :575-636)        */
/* ------------------------------------------------------------------------ */
void dgo_tensor_apply(int dim, const double *Ax, int rx, int cx,
                      const double *Ay, int ry, int cy,
                      const double *Az, int rz, int cz,
                      const double *in, double *out, int adding, double factor)
{
    static const double one = 1.0;
    if (dim == 2) { Az = &one; rz = 1; cz = 1; }
    double *tx = (double *)malloc(sizeof(double) * rx * cy * cz);
    double *txy = (double *)malloc(sizeof(double) * ry * rx * cz);
    /* x pass: tx[x][z][y] */
    for (int x = 0; x < rx; ++x)
        for (int z = 0; z < cz; ++z)
            for (int y = 0; y < cy; ++y) {
                double acc = 0.0;
                for (int s = 0; s < cx; ++s) acc += Ax[x * cx + s] * in[z * cx * cy + y * cx + s];
                tx[x * cy * cz + z * cy + y] = acc;
            }
    /* y pass: txy[y][x][z] */
    for (int y = 0; y < ry; ++y)
        for (int x = 0; x < rx; ++x)
            for (int z = 0; z < cz; ++z) {
                double acc = 0.0;
                for (int s = 0; s < cy; ++s) acc += Ay[y * cy + s] * tx[x * cy * cz + z * cy + s];
                txy[y * rx * cz + x * cz + z] = acc;
            }
    /* z pass: out[z][y][x] */
    for (int z = 0; z < rz; ++z)
        for (int y = 0; y < ry; ++y)
            for (int x = 0; x < rx; ++x) {
                const int idx = z * rx * ry + y * rx + x;
                double o = adding ? out[idx] : 0.0;
                for (int s = 0; s < cz; ++s) o += factor * Az[z * cz + s] * txy[y * rx * cz + x * cz + s];
                out[idx] = o;
            }
    free(tx); free(txy);
}

static void transpose(const double *A, int r, int c, double *At)
{
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < c; ++j) At[j * r + i] = A[i * c + j];
}

/* inner product: out = (Az^T x Ay^T x Ax^T)(in .* w) (operators.cpp:575-636) */
static void inner3(int dim, const double *in, const double *w, double *out,
                   const double *Ax, int rx, int cx, const double *Ay, int ry, int cy,
                   const double *Az, int rz, int cz, int n_in, int adding, double factor)
{
    double Axt[MAXN * MAXN], Ayt[MAXN * MAXN], Azt[MAXN * MAXN];
    transpose(Ax, rx, cx, Axt);
    transpose(Ay, ry, cy, Ayt);
    if (dim == 3) transpose(Az, rz, cz, Azt);
    double *tmp = (double *)malloc(sizeof(double) * n_in);
    for (int i = 0; i < n_in; ++i) tmp[i] = in[i] * w[i];
    dgo_tensor_apply(dim, Axt, cx, rx, Ayt, cy, ry, Azt, cz, rz, tmp, out, adding, factor);
    free(tmp);
}

/* ------------------------------------------------------------------------ */
/* context helpers                                                          */
/* ------------------------------------------------------------------------ */
#define IDX(c, cell, s, node) (((long)(cell) * (c)->ns + (s)) * (c)->Np + (node))

static void vol_apply(const dgo_ctx *c, const double *A, int r, int cc, const double *in, double *out,
                      int adding, double factor)
{
    dgo_tensor_apply(c->dim, A, r, cc, A, r, cc, A, r, cc, in, out, adding, factor);
}

/* apply with operator G in direction dir and V in the others */
static void dir_apply(const dgo_ctx *c, int dir, const double *G, const double *V, int r, int cc,
                      const double *in, double *out, int adding, double factor)
{
    dgo_tensor_apply(c->dim, dir == 0 ? G : V, r, cc, dir == 1 ? G : V, r, cc, dir == 2 ? G : V, r, cc,
                     in, out, adding, factor);
}

/* matrix_vector_mult_surface_1D (operators.cpp:414-440): surf row S (1 x cc) in the face's
   normal direction, V (r x cc) tangentially */
static void surf_apply(const dgo_ctx *c, int face, const double *S, const double *V, int r, int cc,
                       const double *in, double *out, int adding, double factor)
{
    const int dir = face / 2;
    dgo_tensor_apply(c->dim, dir == 0 ? S : V, dir == 0 ? 1 : r, cc,
                     dir == 1 ? S : V, dir == 1 ? 1 : r, cc,
                     dir == 2 ? S : V, dir == 2 ? 1 : r, cc, in, out, adding, factor);
}

/* inner_product_surface_1D (operators.cpp:445-477) */
static void surf_inner(const dgo_ctx *c, int face, const double *in, const double *w, double *out,
                       const double *S, const double *V, int r, int cc, int adding, double factor)
{
    const int dir = face / 2;
    inner3(c->dim, in, w, out,
           dir == 0 ? S : V, dir == 0 ? 1 : r, cc,
           dir == 1 ? S : V, dir == 1 ? 1 : r, cc,
           dir == 2 ? S : V, dir == 2 ? 1 : r, cc, c->nfp, adding, factor);
}

static void vol_inner(const dgo_ctx *c, const double *in, const double *w, double *out, int adding, double factor)
{
    inner3(c->dim, in, w, out, c->chi, c->nq, c->n, c->chi, c->nq, c->n, c->chi, c->nq, c->n, c->Np,
           adding, factor);
}

/* ------------------------------------------------------------------------ */
/* Hadamard sparsity (operators.cpp:949-1012, :1067-1130)                   */
/* ------------------------------------------------------------------------ */
void dgo_hadamard_volume_pattern(int dim, int n, unsigned *rows, unsigned *cols)
{
    if (dim == 2) {
        for (unsigned i = 0; i < (unsigned)n; ++i)
            for (unsigned j = 0; j < (unsigned)n; ++j)
                for (unsigned k = 0; k < (unsigned)n; ++k) {
                    const unsigned a = i * n * n + j * n + k;
                    const unsigned row = i * n;
                    rows[a * 2 + 0] = row + j;
                    rows[a * 2 + 1] = row + j;
                    cols[a * 2 + 0] = i * n + k;
                    cols[a * 2 + 1] = k * n + j;
                }
    } else {
        for (unsigned i = 0; i < (unsigned)n; ++i)
            for (unsigned j = 0; j < (unsigned)n; ++j)
                for (unsigned k = 0; k < (unsigned)n; ++k)
                    for (unsigned l = 0; l < (unsigned)n; ++l) {
                        const unsigned a = i * n * n * n + j * n * n + k * n + l;
                        const unsigned row = i * n * n + j * n;
                        rows[a * 3 + 0] = row + k;
                        rows[a * 3 + 1] = row + k;
                        rows[a * 3 + 2] = row + k;
                        cols[a * 3 + 0] = i * n * n + j * n + l;
                        cols[a * 3 + 1] = l * n + k + i * n * n;
                        cols[a * 3 + 2] = l * n * n + k + j * n;
                    }
    }
}

void dgo_hadamard_surface_pattern(int dim, int n, int dnz, unsigned *rows, unsigned *cols)
{
    if (dim == 2) {
        for (unsigned i = 0; i < (unsigned)n; ++i)
            for (unsigned j = 0; j < (unsigned)n; ++j) {
                const unsigned a = i * n + j;
                rows[a] = i;
                cols[a] = (dnz == 0) ? i * n + j : j * n + i;
            }
    } else {
        for (unsigned i = 0; i < (unsigned)n; ++i)
            for (unsigned j = 0; j < (unsigned)n; ++j)
                for (unsigned k = 0; k < (unsigned)n; ++k) {
                    const unsigned a = i * n * n + j * n + k;
                    rows[a] = i * n + j;
                    if (dnz == 0) cols[a] = i * n * n + j * n + k;
                    if (dnz == 1) cols[a] = k * n + j + i * n * n;
                    if (dnz == 2) cols[a] = k * n * n + j + i * n;
                }
    }
}

/* ------------------------------------------------------------------------ */
/* mesh: straight_periodic_cube (src/mesh/grids/straight_periodic_cube.cpp) */
/* ------------------------------------------------------------------------ */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void build_mesh(dgo_ctx *c)
{
    const int N = c->P.n_cells_1d, d = c->dim;
    c->cell_ijk = (int *)calloc((size_t)c->ncells * 3, sizeof(int));
    c->nbr = (int *)calloc((size_t)c->ncells * 2 * d, sizeof(int));
    int *lookup = (int *)malloc(sizeof(int) * c->ncells);
    const int morton = (c->P.cell_order == DGO_ORDER_DEALII) && is_pow2(N);
    int levels = 0;
    while ((1 << levels) < N) ++levels;
    for (int e = 0; e < c->ncells; ++e) {
        int ijk[3] = {0, 0, 0};
        if (morton) {
            /* refine_global: child index bits x,y(,z) per level, coarsest first */
            for (int l = 0; l < levels; ++l) {
                const int shift = (levels - 1 - l);
                const int child = (e >> (d * shift)) & ((1 << d) - 1);
                for (int k = 0; k < d; ++k) ijk[k] |= ((child >> k) & 1) << shift;
            }
        } else {
            ijk[0] = e % N;
            ijk[1] = (e / N) % N;
            if (d == 3) ijk[2] = e / (N * N);
        }
        for (int k = 0; k < 3; ++k) c->cell_ijk[e * 3 + k] = ijk[k];
        const int lin = ijk[0] + N * (ijk[1] + (d == 3 ? N * ijk[2] : 0));
        lookup[lin] = e;
    }
    for (int e = 0; e < c->ncells; ++e) {
        for (int f = 0; f < 2 * d; ++f) {
            int ijk[3] = {c->cell_ijk[e * 3], c->cell_ijk[e * 3 + 1], c->cell_ijk[e * 3 + 2]};
            const int k = f / 2;
            /* a face on the box boundary of a non-periodic direction: -(boundary type) */
            if (c->P.boundary_type[f] && ijk[k] == (f % 2 ? N - 1 : 0)) {
                c->nbr[e * 2 * d + f] = -c->P.boundary_type[f];
                continue;
            }
            ijk[k] = (ijk[k] + (f % 2 ? 1 : N - 1)) % N;
            const int lin = ijk[0] + N * (ijk[1] + (d == 3 ? N * ijk[2] : 0));
            c->nbr[e * 2 * d + f] = lookup[lin];
        }
    }
    free(lookup);
}

void dgo_cell_ijk(const dgo_ctx *c, int *ijk) { memcpy(ijk, c->cell_ijk, sizeof(int) * 3 * c->ncells); }
void dgo_neighbors(const dgo_ctx *c, int *nbr) { memcpy(nbr, c->nbr, sizeof(int) * 2 * c->dim * c->ncells); }

/* ------------------------------------------------------------------------ */
/* boundary states: Euler::boundary_face_values (euler.cpp:1511-1567)        */
/* ------------------------------------------------------------------------ */
static void prim_of(const dgo_ctx *c, const double *u, double *w)
{
    /* convert_conservative_to_primitive: rho, velocity, pressure */
    const int d = c->dim;
    double v2 = 0.0;
    w[0] = u[0];
    for (int k = 0; k < d; ++k) {
        w[1 + k] = u[1 + k] / u[0];
        v2 += w[1 + k] * w[1 + k];
    }
    w[d + 1] = c->gamm1 * (u[d + 1] - 0.5 * u[0] * v2);
}

static void cons_of(const dgo_ctx *c, const double *w, double *u)
{
    /* convert_primitive_to_conservative */
    const int d = c->dim;
    double v2 = 0.0;
    u[0] = w[0];
    for (int k = 0; k < d; ++k) {
        u[1 + k] = w[0] * w[1 + k];
        v2 += w[1 + k] * w[1 + k];
    }
    u[d + 1] = w[d + 1] / c->gamm1 + 0.5 * w[0] * v2;
}

void dgo_boundary_state(const dgo_ctx *c, int type, const double *u_int, const double *n, double *u_bc)
{
    const int d = c->dim, ns = c->ns;
    const double gam = c->gam, gamm1 = c->gamm1, pinf = c->pinf, minf2 = c->mach2;
    /* total inlet pressure and temperature (euler.cpp:1518-1519); temperature_inf = 1 */
    const double p0 = pinf * pow(1.0 + 0.5 * gamm1 * minf2, gam / gamm1);
    const double T0 = 1.0 * pow(p0 / pinf, gamm1 / gam);
    double wi[NS_MAX], wb[NS_MAX];
    prim_of(c, u_int, wi);
    if (type == 1001 || type == 1006) {
        /* slip wall (euler.cpp:1171-1215; 1001 is the same for Euler, :1219-1228): the normal
           velocity reflected, density and pressure copied */
        double vn = 0.0;
        for (int k = 0; k < d; ++k) vn += wi[1 + k] * (-n[k]);
        wb[0] = wi[0];
        for (int k = 0; k < d; ++k) wb[1 + k] = wi[1 + k] - 2.0 * vn * (-n[k]);
        wb[d + 1] = wi[d + 1];
        cons_of(c, wb, u_bc);
    } else if (type == 1002) {
        /* pressure outflow, back pressure 0.99 of the total inlet pressure (:1290-1325) */
        double v2 = 0.0;
        for (int k = 0; k < d; ++k) v2 += wi[1 + k] * wi[1 + k];
        const double mach_int = sqrt(v2) / sqrt(gam * wi[d + 1] / wi[0]);
        if (mach_int > 1.0) {
            for (int s = 0; s < ns; ++s) u_bc[s] = u_int[s];
            return;
        }
        const double p_inlet = p0 * pow(1.0 + 0.5 * gamm1 * minf2, -gam / gamm1);
        const double p_bc = (mach_int >= 1.0) ? wi[d + 1] : 0.99 * p_inlet;
        const double T_int = gam * minf2 * (wi[d + 1] / wi[0]);
        wb[0] = gam * minf2 * (p_bc / T_int);
        for (int k = 0; k < d; ++k) wb[1 + k] = wi[1 + k];
        wb[d + 1] = p_bc;
        cons_of(c, wb, u_bc);
    } else if (type == 1003) {
        /* inflow (:1327-1431, Carlson 2011 sec. 2.7 / 2.9) */
        double vn = 0.0;
        for (int k = 0; k < d; ++k) vn += wi[1 + k] * (-n[k]);
        const double snd = sqrt(gam * wi[d + 1] / wi[0]);
        if (minf2 < 1.0) {
            const double rp = vn + 2.0 * snd / gamm1;
            const double H = u_int[d + 1] / wi[0] + wi[d + 1] / wi[0];
            const double A = 1.0 + 2.0 / gamm1, B = -2.0 * rp, C = 0.5 * gamm1 * (rp * rp - 2.0 * H);
            const double disc = sqrt(B * B - 4.0 * A * C);
            const double c1 = (-B + disc) / (2.0 * A), c2 = (-B - disc) / (2.0 * A);
            const double cb = c1 > c2 ? c1 : c2;
            const double vb = rp - 2.0 * cb / gamm1;
            const double mb = vb / cb;
            const double rad = 1.0 + 0.5 * gamm1 * mb * mb;
            const double pb = p0 * pow(rad, -gam / gamm1), Tb = T0 / rad;
            wb[0] = gam * minf2 * (pb / Tb);
            for (int k = 0; k < d; ++k) wb[1 + k] = vb * (-n[k]);
            wb[d + 1] = pb;
        } else {
            const double rad = 1.0 + 0.5 * gamm1 * minf2;
            const double pb = p0 * pow(rad, -gam / gamm1), Tb = T0 / rad;
            wb[0] = gam * minf2 * (pb / Tb);
            const double vmag = sqrt(minf2) * sqrt(gam * pb / wb[0]);
            for (int k = 0; k < d; ++k) wb[1 + k] = -vmag * n[k];
            wb[d + 1] = pb;
        }
        cons_of(c, wb, u_bc);
    } else if (type == 1004) {
        /* Riemann-invariant farfield (:1097-1169) */
        double we[NS_MAX];
        we[0] = 1.0;
        for (int k = 0; k < d; ++k) we[1 + k] = c->vinf[k];
        we[d + 1] = pinf;
        const double ci = sqrt(gam * wi[d + 1] / wi[0]), ce = sqrt(gam * we[d + 1] / we[0]);
        double vni = 0.0, vne = 0.0;
        for (int k = 0; k < d; ++k) {
            vni += wi[1 + k] * n[k];
            vne += we[1 + k] * n[k];
        }
        const double Rout = vni + 2.0 / gamm1 * ci, Rinc = vne - 2.0 / gamm1 * ce;
        const double vnb = 0.5 * (Rout + Rinc), cb = 0.25 * gamm1 * (Rout - Rinc);
        if (fabs(vnb) >= fabs(cb)) {
            for (int s = 0; s < ns; ++s) wb[s] = (vnb < 0.0) ? we[s] : wi[s];
        } else {
            const double *w_src = (vnb < 0.0) ? we : wi;
            const double vn_src = (vnb < 0.0) ? vne : vni;
            const double ent = w_src[d + 1] * pow(w_src[0], -gam);  /* compute_entropy_measure */
            wb[0] = pow(1.0 / gam * cb * cb / ent, 1.0 / gamm1);
            for (int k = 0; k < d; ++k) wb[1 + k] = (w_src[1 + k] - vn_src * n[k]) + vnb * n[k];
            wb[d + 1] = 1.0 / gam * cb * cb * wb[0];
        }
        cons_of(c, wb, u_bc);
    } else if (type == 1005) {
        /* free stream (:1433-1447) */
        wb[0] = 1.0;
        for (int k = 0; k < d; ++k) wb[1 + k] = c->vinf[k];
        wb[d + 1] = 1.0 / (gam * minf2);
        cons_of(c, wb, u_bc);
    } else {
        /* 1007: extrapolation (:1449-1460) */
        for (int s = 0; s < ns; ++s) u_bc[s] = u_int[s];
    }
}

/* NavierStokes::boundary_face_values_viscous_flux for the wall types (navier_stokes.cpp:1354-1430):
   1001 no-slip wall -- momentum reversed, the gradient copied with the energy row zeroed on an
   adiabatic wall, or the energy set so that the average temperature is the wall temperature on an
   isothermal one; 1006 slip wall -- the interior state and gradient.  g/gb: [s*dim + k]. */
static void visc_boundary_state(const dgo_ctx *c, int type, const double *ui, const double *gi, double *ub,
                                double *gb)
{
    const int ns = c->ns, d = c->dim;
    for (int s = 0; s < ns; ++s) {
        ub[s] = ui[s];
        for (int k = 0; k < d; ++k) gb[s * d + k] = gi[s * d + k];
    }
    if (type != 1001) return;
    for (int k = 0; k < d; ++k) ub[1 + k] = -ui[1 + k];
    if (c->P.thermal_bc == 0) {
        for (int k = 0; k < d; ++k) gb[(ns - 1) * d + k] = 0.0;
    } else {
        double wi[NS_MAX], we[NS_MAX];
        prim_of(c, ui, wi);
        prim_of(c, ub, we);
        const double T_int = c->gam * c->mach2 * (wi[ns - 1] / wi[0]);
        const double T_ext = 2.0 * c->P.wall_temperature - T_int;
        we[ns - 1] = we[0] * T_ext / (c->gam * c->mach2);
        double v2 = 0.0;
        for (int k = 0; k < d; ++k) v2 += we[1 + k] * we[1 + k];
        ub[ns - 1] = we[ns - 1] / c->gamm1 + 0.5 * we[0] * v2;
    }
}

static int wall_adiabatic(const dgo_ctx *c, int type) { return type == 1001 && c->P.thermal_bc == 0; }

/* ------------------------------------------------------------------------ */
/* geometry: node coordinates and metric terms                              */
/* ------------------------------------------------------------------------ */
/* NonsymmetricCurvedGridManifold::mapping (nonsymmetric_curved_periodic_grid.cpp:66-87), beta = 0.1:
   the reference's curved periodic grid on [-1,1]^2 / [0,2pi]^3 */
static void warp_point(int dim, double *x)
{
    const double beta = 1.0 / 10.0, pi = 3.14159265358979323846;
    const double p0 = x[0], p1 = x[1], p2 = x[2];
    if (dim == 2) {
        x[0] = p0 + beta * cos(pi / 2.0 * p0) * cos(3.0 * pi / 2.0 * p1);
        x[1] = p1 + beta * sin(2.0 * pi * p0) * cos(pi / 2.0 * p1);
    } else {
        x[0] = p0 + beta * sin(p0) * sin(p1) * sin(2.0 * p2);
        x[1] = p1 + beta * sin(4.0 * p0) * sin(p1) * sin(3.0 * p2);
        x[2] = p2 + beta * sin(2.0 * p0) * sin(5.0 * p1) * sin(p2);
    }
}

static void build_coords(dgo_ctx *c)
{
    const int n = c->n, Np = c->Np;
    c->xn = (double *)malloc(sizeof(double) * (size_t)c->ncells * Np * 3);
    for (int e = 0; e < c->ncells; ++e)
        for (int node = 0; node < Np; ++node) {
            int id[3] = {node % n, (node / n) % n, (c->dim == 3) ? node / (n * n) : 0};
            double *x = c->xn + ((long)e * Np + node) * 3;
            for (int k = 0; k < 3; ++k)
                x[k] = (k < c->dim) ? c->P.domain_left + (c->cell_ijk[e * 3 + k] + c->xs[id[k]]) * c->h : 0.0;
            if (c->P.grid_warp) warp_point(c->dim, x);
        }
}

/* Metric terms at the GLL flux nodes from the mapping support points at the same nodes (grid
   degree = p): the Jacobian dX_i/dxi_j by the 1D derivative operator (build_metric_Jacobian,
   operators.cpp:2600-2650), det J, and the cofactor C[m][n] = J dxi_n/dx_m -- in 2D from the
   Jacobian (operators.cpp:2700-2714), in 3D by the conservative curl form (compute_local_3D_cofactor,
   operators.cpp:2737-2884; use_invariant_curl_form = false), which satisfies the discrete metric
   identities sum_n D_n C[m][n] = 0. */
static void compute_metrics(dgo_ctx *c)
{
    const int d = c->dim, n = c->n, Np = c->Np;
    free(c->metC); free(c->metJ);
    c->metC = (double *)calloc((size_t)c->ncells * Np * 9, sizeof(double));
    c->metJ = (double *)malloc(sizeof(double) * (size_t)c->ncells * Np);
    double *X = (double *)malloc(sizeof(double) * 3 * Np);
    double *Jm = (double *)malloc(sizeof(double) * 9 * Np);   /* Jm[(i*3+j)*Np + q] = dX_i/dxi_j */
    double *P = (double *)malloc(sizeof(double) * 9 * Np);    /* curl-form products */
    double *Cm = (double *)malloc(sizeof(double) * 9 * Np);
    for (int e = 0; e < c->ncells; ++e) {
        for (int i = 0; i < d; ++i)
            for (int q = 0; q < Np; ++q) X[i * Np + q] = c->xn[((long)e * Np + q) * 3 + i];
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) dir_apply(c, j, c->chi_grad, c->chi, n, n, X + i * Np, Jm + (i * 3 + j) * Np, 0, 1.0);
        if (d == 2) {
            for (int q = 0; q < Np; ++q) {
                Cm[(0 * 3 + 0) * Np + q] = Jm[(1 * 3 + 1) * Np + q];
                Cm[(1 * 3 + 0) * Np + q] = -Jm[(0 * 3 + 1) * Np + q];
                Cm[(0 * 3 + 1) * Np + q] = -Jm[(1 * 3 + 0) * Np + q];
                Cm[(1 * 3 + 1) * Np + q] = Jm[(0 * 3 + 0) * Np + q];
            }
        } else {
            /* X_l dX_m/dxi_j at the grid nodes: z dy/d(xi,eta,zeta), x dz/d(..), y dx/d(..) */
            const int lm[3][2] = {{2, 1}, {0, 2}, {1, 0}};
            for (int r = 0; r < 3; ++r)
                for (int j = 0; j < 3; ++j)
                    for (int q = 0; q < Np; ++q)
                        P[(r * 3 + j) * Np + q] = X[lm[r][0] * Np + q] * Jm[(lm[r][1] * 3 + j) * Np + q];
            for (int r = 0; r < 3; ++r) {
                const double *Pxi = P + (r * 3 + 0) * Np, *Peta = P + (r * 3 + 1) * Np, *Pzeta = P + (r * 3 + 2) * Np;
                double *C0 = Cm + (r * 3 + 0) * Np, *C1 = Cm + (r * 3 + 1) * Np, *C2 = Cm + (r * 3 + 2) * Np;
                /* C[r][0] = -d_eta(P_zeta) + d_zeta(P_eta); C[r][1] = d_xi(P_zeta) - d_zeta(P_xi);
                   C[r][2] = -d_xi(P_eta) + d_eta(P_xi) */
                dir_apply(c, 1, c->chi_grad, c->chi, n, n, Pzeta, C0, 0, -1.0);
                dir_apply(c, 2, c->chi_grad, c->chi, n, n, Peta, C0, 1, 1.0);
                dir_apply(c, 0, c->chi_grad, c->chi, n, n, Pzeta, C1, 0, 1.0);
                dir_apply(c, 2, c->chi_grad, c->chi, n, n, Pxi, C1, 1, -1.0);
                dir_apply(c, 0, c->chi_grad, c->chi, n, n, Peta, C2, 0, -1.0);
                dir_apply(c, 1, c->chi_grad, c->chi, n, n, Pxi, C2, 1, 1.0);
            }
        }
        for (int q = 0; q < Np; ++q) {
            double J[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) J[i][j] = (i < d && j < d) ? Jm[(i * 3 + j) * Np + q] : 0.0;
            double det;
            if (d == 2) det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            else det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                     + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            c->metJ[(long)e * Np + q] = det;
            if (d == 3 && g_mutation == DGO_MUT_CROSS_PRODUCT_METRIC)
                /* planted: the pointwise cofactor det(J) J^-T (cross products of the Jacobian's columns) */
                for (int m = 0; m < 3; ++m)
                    for (int nn = 0; nn < 3; ++nn) {
                        const int m1 = (m + 1) % 3, m2 = (m + 2) % 3, n1 = (nn + 1) % 3, n2 = (nn + 2) % 3;
                        Cm[(m * 3 + nn) * Np + q] = J[m1][n1] * J[m2][n2] - J[m1][n2] * J[m2][n1];
                    }
            for (int a = 0; a < d; ++a)
                for (int b = 0; b < d; ++b) c->metC[((long)e * Np + q) * 9 + a * 3 + b] = Cm[(a * 3 + b) * Np + q];
        }
    }
    free(X); free(Jm); free(P); free(Cm);
}

int dgo_set_grid_nodes(dgo_ctx *c, const double *xyz)
{
    if (c->P.flux_nodes != DGO_GLL) return -1;
    memcpy(c->xn, xyz, sizeof(double) * (size_t)c->ncells * c->Np * 3);
    compute_metrics(c);
    c->curv = 1;
    return 0;
}

void dgo_metric_terms(const dgo_ctx *c, double *C, double *J)
{
    const long nn = (long)c->ncells * c->Np;
    for (long q = 0; q < nn; ++q) {
        const int e = (int)(q / c->Np), node = (int)(q % c->Np);
        for (int a = 0; a < 9; ++a) {
            double v = 0.0;
            if (c->curv) v = c->metC[q * 9 + a];
            else if (a / 3 == a % 3 && a / 3 < c->dim) v = c->cof;
            C[q * 9 + a] = v;
        }
        J[q] = c->curv ? c->metJ[(long)e * c->Np + node] : c->detJ;
    }
}

/* owner rule current_cell_should_do_the_work (dg_base.cpp:417-449): lower index works */
int dgo_n_faces(const dgo_ctx *c)
{
    int nf = 0;
    for (int e = 0; e < c->ncells; ++e)
        for (int f = 0; f < 2 * c->dim; ++f)
            if (e < c->nbr[e * 2 * c->dim + f]) ++nf;
    return nf;
}

void dgo_face_list(const dgo_ctx *c, int *faces)
{
    int nf = 0;
    for (int e = 0; e < c->ncells; ++e)
        for (int f = 0; f < 2 * c->dim; ++f) {
            const int nb = c->nbr[e * 2 * c->dim + f];
            if (e < nb) {
                faces[nf * 4 + 0] = e;
                faces[nf * 4 + 1] = f;
                faces[nf * 4 + 2] = nb;
                faces[nf * 4 + 3] = f ^ 1;
                ++nf;
            }
        }
}

void dgo_node_coords(const dgo_ctx *c, double *xyz)
{
    memcpy(xyz, c->xn, sizeof(double) * (size_t)c->ncells * c->Np * 3);
}

/* ------------------------------------------------------------------------ */
/* context                                                                   */
/* ------------------------------------------------------------------------ */
void dgo_default_params(dgo_params *p)
{
    memset(p, 0, sizeof(*p));
    p->dim = 3;
    p->poly_degree = 2;
    p->flux_nodes = DGO_GLL;
    p->use_split_form = 0;
    p->two_point_flux = DGO_KG;
    p->conv_flux = DGO_LXF;
    p->pde = DGO_EULER;
    p->diss_flux = DGO_SIPG;
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
    p->cell_order = DGO_ORDER_DEALII;
    for (int f = 0; f < 6; ++f) p->boundary_type[f] = 0;
    p->angle_of_attack = 0.0;
    p->grid_warp = 0;
    p->thermal_bc = 0;
    p->wall_temperature = 1.0;
}

dgo_ctx *dgo_create(const dgo_params *prm)
{
    dgo_ctx *c = (dgo_ctx *)calloc(1, sizeof(dgo_ctx));
    c->P = *prm;
    c->dim = prm->dim;
    c->p = prm->poly_degree;
    c->n = c->p + 1;
    c->nq = c->n; /* overintegration = 0 */
    c->Np = (c->dim == 3) ? c->n * c->n * c->n : c->n * c->n;
    c->nfp = (c->dim == 3) ? c->n * c->n : c->n;
    c->ns = c->dim + 2;
    c->ncells = (c->dim == 3) ? prm->n_cells_1d * prm->n_cells_1d * prm->n_cells_1d
                              : prm->n_cells_1d * prm->n_cells_1d;
    c->h = (prm->domain_right - prm->domain_left) / prm->n_cells_1d;
    c->detJ = (c->dim == 3) ? c->h * c->h * c->h : c->h * c->h;
    c->cof = (c->dim == 3) ? c->h * c->h : c->h;
    c->one1[0] = 1.0;
    c->fsal = 1;

    c->gam = prm->gamma_gas;
    c->gamm1 = c->gam - 1.0;
    c->mach2 = prm->mach_inf * prm->mach_inf;
    c->pinf = 1.0 / (c->gam * c->mach2);
    c->temperature_ratio = 110.4 / prm->temperature_inf;
    c->vinf[0] = cos(prm->angle_of_attack);
    c->vinf[1] = sin(prm->angle_of_attack);
    c->vinf[2] = 0.0;
    /* boundaries: Euler only, both faces of a direction periodic or both boundaries, known types */
    for (int f = 0; f < 2 * c->dim; ++f) {
        const int t = prm->boundary_type[f], o = prm->boundary_type[f ^ 1];
        /* Navier-Stokes: the wall types only (boundary_face_values_viscous_flux sets the viscous
           boundary state for 1001 and 1006 and leaves it unset for the others) */
        const int ns_ok = prm->pde == DGO_EULER || t == 1001 || t == 1006;
        if ((t == 0) != (o == 0) || (t && (t < 1001 || t > 1007)) || (t && !ns_ok) ||
            (prm->thermal_bc != 0 && prm->thermal_bc != 1) || !(prm->wall_temperature > 0.0)) {
            free(c);
            return NULL;
        }
    }
    /* curvilinear grids: collocated GLL nodes (grid degree = p) */
    if (prm->grid_warp && (prm->flux_nodes != DGO_GLL || prm->grid_warp != 1)) {
        free(c);
        return NULL;
    }

    const int n = c->n, nq = c->nq;
    /* FE_DGQ(p): Lagrange on Gauss-Lobatto support points (dg_base.cpp:209-217) */
    dgo_gauss_lobatto(n, c->xs, c->ws);
    if (prm->flux_nodes == DGO_GLL) dgo_gauss_lobatto(nq, c->xq, c->wq);
    else dgo_gauss_legendre(nq, c->xq, c->wq);

    double val[MAXN], der[MAXN];
    for (int q = 0; q < nq; ++q) {
        dgo_lagrange_values(n, c->xs, c->xq[q], val, der);
        for (int j = 0; j < n; ++j) { c->chi[q * n + j] = val[j]; c->chi_grad[q * n + j] = der[j]; }
        dgo_lagrange_values(nq, c->xq, c->xq[q], val, der);
        for (int j = 0; j < nq; ++j) { c->fb_vol[q * nq + j] = val[j]; c->fb_grad[q * nq + j] = der[j]; }
    }
    for (int f = 0; f < 2; ++f) {
        dgo_lagrange_values(n, c->xs, (double)f, val, NULL);
        for (int j = 0; j < n; ++j) c->chi_surf[f][j] = val[j];
        dgo_lagrange_values(nq, c->xq, (double)f, val, NULL);
        for (int j = 0; j < nq; ++j) c->fb_surf[f][j] = val[j];
    }
    /* local_basis_stiffness on the flux basis (operators.cpp:1440-1477) */
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nq; ++j) {
            double v = 0.0;
            for (int q = 0; q < nq; ++q) v += c->fb_vol[q * nq + i] * c->fb_grad[q * nq + j] * c->wq[q];
            c->stiff[i * nq + j] = v;
        }
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nq; ++j) c->skew[i * nq + j] = c->stiff[i * nq + j] - c->stiff[j * nq + i];
    /* local_mass on the soln basis (operators.cpp:1354-1385), cDG => FR_mass_inv = M^-1 */
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double v = 0.0;
            for (int q = 0; q < nq; ++q) v += c->chi[q * n + i] * c->chi[q * n + j] * c->wq[q];
            c->mass[i * n + j] = v;
        }
    invert(n, c->mass, c->minv);
    /* vol_projection_operator = M^-1 (W chi)^T (operators.cpp:1873-1889) */
    for (int i = 0; i < n; ++i)
        for (int q = 0; q < nq; ++q) {
            double v = 0.0;
            for (int j = 0; j < n; ++j) v += c->minv[i * n + j] * c->wq[q] * c->chi[q * n + j];
            c->proj[i * nq + q] = v;
        }

    const int Np = c->Np, nfp = c->nfp;
    c->wvol = (double *)malloc(sizeof(double) * Np);
    c->wface = (double *)malloc(sizeof(double) * nfp);
    c->ones_vol = (double *)malloc(sizeof(double) * Np);
    c->ones_face = (double *)malloc(sizeof(double) * nfp);
    for (int node = 0; node < Np; ++node) {
        const int ix = node % n, iy = (node / n) % n, iz = (c->dim == 3) ? node / (n * n) : 0;
        c->wvol[node] = (c->dim == 3) ? c->wq[ix] * c->wq[iy] * c->wq[iz] : c->wq[ix] * c->wq[iy];
        c->ones_vol[node] = 1.0;
    }
    for (int fnode = 0; fnode < nfp; ++fnode) {
        const int a = fnode % n, b = fnode / n;
        c->wface[fnode] = (c->dim == 3) ? c->wq[a] * c->wq[b] : c->wq[a];
        c->ones_face[fnode] = 1.0;
    }
    c->had_rows = (unsigned *)malloc(sizeof(unsigned) * Np * nq * c->dim);
    c->had_cols = (unsigned *)malloc(sizeof(unsigned) * Np * nq * c->dim);
    dgo_hadamard_volume_pattern(c->dim, nq, c->had_rows, c->had_cols);
    for (int k = 0; k < c->dim; ++k) {
        c->had_srows[k] = (unsigned *)malloc(sizeof(unsigned) * nfp * nq);
        c->had_scols[k] = (unsigned *)malloc(sizeof(unsigned) * nfp * nq);
        dgo_hadamard_surface_pattern(c->dim, nq, k, c->had_srows[k], c->had_scols[k]);
    }
    build_mesh(c);
    build_coords(c);
    if (prm->grid_warp) {
        compute_metrics(c);
        c->curv = 1;
    }
    if (prm->pde == DGO_NAVIER_STOKES) {
        for (int k = 0; k < c->dim; ++k) {
            c->aux[k] = (double *)calloc((size_t)dgo_n_dofs(c), sizeof(double));
            c->aux_rhs[k] = (double *)calloc((size_t)dgo_n_dofs(c), sizeof(double));
        }
    }
    return c;
}

void dgo_destroy(dgo_ctx *c)
{
    if (!c) return;
    free(c->wvol); free(c->wface); free(c->ones_vol); free(c->ones_face);
    free(c->had_rows); free(c->had_cols);
    for (int k = 0; k < 3; ++k) {
        free(c->had_srows[k]); free(c->had_scols[k]);
        free(c->aux[k]); free(c->aux_rhs[k]);
    }
    free(c->cell_ijk); free(c->nbr);
    free(c->xn); free(c->metC); free(c->metJ);
    free(c);
}

long dgo_n_dofs(const dgo_ctx *c) { return (long)c->ncells * c->ns * c->Np; }
int dgo_n_cells(const dgo_ctx *c) { return c->ncells; }

void dgo_get_1d_operators(const dgo_ctx *c, double *nodes, double *weights,
                          double *chi, double *chi_grad, double *flux_grad,
                          double *skew, double *proj, double *minv,
                          double *chi_surf, double *flux_surf)
{
    const int n = c->n, nq = c->nq;
    if (nodes) memcpy(nodes, c->xq, sizeof(double) * nq);
    if (weights) memcpy(weights, c->wq, sizeof(double) * nq);
    if (chi) memcpy(chi, c->chi, sizeof(double) * nq * n);
    if (chi_grad) memcpy(chi_grad, c->chi_grad, sizeof(double) * nq * n);
    if (flux_grad) memcpy(flux_grad, c->fb_grad, sizeof(double) * nq * nq);
    if (skew) memcpy(skew, c->skew, sizeof(double) * nq * nq);
    if (proj) memcpy(proj, c->proj, sizeof(double) * n * nq);
    if (minv) memcpy(minv, c->minv, sizeof(double) * n * n);
    if (chi_surf) { memcpy(chi_surf, c->chi_surf[0], sizeof(double) * n); memcpy(chi_surf + n, c->chi_surf[1], sizeof(double) * n); }
    if (flux_surf) { memcpy(flux_surf, c->fb_surf[0], sizeof(double) * nq); memcpy(flux_surf + nq, c->fb_surf[1], sizeof(double) * nq); }
}

/* ------------------------------------------------------------------------ */
/* Euler physics (src/physics/euler.cpp)                                    */
/* ------------------------------------------------------------------------ */
double dgo_pressure(const dgo_ctx *c, const double *u)
{
    /* compute_pressure_templated (euler.cpp:486-500) */
    const int d = c->dim;
    const double rho = u[0], E = u[d + 1];
    double vel2 = 0.0;
    for (int k = 0; k < d; ++k) { const double v = u[1 + k] / rho; vel2 = vel2 + v * v; }
    return c->gamm1 * (E - 0.5 * rho * vel2);
}

static void velocities(const dgo_ctx *c, const double *u, double *vel)
{
    for (int k = 0; k < c->dim; ++k) vel[k] = u[1 + k] / u[0];
}

void dgo_convective_flux(const dgo_ctx *c, const double *u, double *f)
{
    /* euler.cpp:915-937 ; f[s*dim + flux_dim] */
    const int d = c->dim, ns = c->ns;
    const double rho = u[0];
    const double p = dgo_pressure(c, u);
    double vel[3];
    velocities(c, u, vel);
    const double e = u[ns - 1] / u[0];
    const double H = e + p / rho;
    for (int fd = 0; fd < d; ++fd) {
        f[0 * d + fd] = u[1 + fd];
        for (int vd = 0; vd < d; ++vd) f[(1 + vd) * d + fd] = rho * vel[fd] * vel[vd];
        f[(1 + fd) * d + fd] += p;
        f[(ns - 1) * d + fd] = rho * vel[fd] * H;
    }
}

void dgo_entropy_variables(const dgo_ctx *c, const double *u, double *v)
{
    /* compute_entropy_variables (euler.cpp:810-829), entropy = log(p rho^-gamma) (:511-532) */
    const int d = c->dim, ns = c->ns;
    const double rho = u[0];
    const double p = dgo_pressure(c, u);
    const double s = log(p * pow(rho, -c->gam));
    const double rho_theta = p / c->gamm1;
    v[0] = (rho_theta * (c->gam + 1.0 - s) - u[ns - 1]) / rho_theta;
    for (int k = 0; k < d; ++k) v[k + 1] = u[k + 1] / rho_theta;
    v[ns - 1] = -rho / rho_theta;
}

void dgo_conservative_from_entropy(const dgo_ctx *c, const double *v, double *u)
{
    /* euler.cpp:832-853 */
    const int d = c->dim, ns = c->ns;
    double vv = 0.0;
    for (int k = 0; k < d; ++k) vv += v[k + 1] * v[k + 1];
    const double s = c->gam - v[0] + 0.5 * vv / v[ns - 1];
    const double rho_theta = pow((c->gamm1 / pow(-v[ns - 1], c->gam)), 1.0 / c->gamm1) * exp(-s / c->gamm1);
    u[0] = -rho_theta * v[ns - 1];
    for (int k = 0; k < d; ++k) u[k + 1] = rho_theta * v[k + 1];
    u[ns - 1] = rho_theta * (1.0 - 0.5 * vv / v[ns - 1]);
}

double dgo_log_mean(double a, double b)
{
    /* compute_ismail_roe_logarithmic_mean (euler.cpp:629-647) */
    const double zeta = a / b;
    const double f = (zeta - 1.0) / (zeta + 1.0);
    const double uu = f * f;
    double F;
    if (uu < 1.0e-2) F = 1.0 + uu / 3.0 + uu * uu / 5.0 + uu * uu * uu / 7.0;
    else F = log(zeta) / 2.0 / f;
    return (a + b) / (2.0 * F);
}

static void flux_kg(const dgo_ctx *c, const double *u1, const double *u2, double *f)
{
    /* euler.cpp:587-611 */
    const int d = c->dim, ns = c->ns;
    const double mrho = (u1[0] + u2[0]) / 2.;
    const double mp = (dgo_pressure(c, u1) + dgo_pressure(c, u2)) / 2.;
    double v1[3], v2[3], mv[3];
    velocities(c, u1, v1);
    velocities(c, u2, v2);
    for (int k = 0; k < d; ++k) mv[k] = 0.5 * (v1[k] + v2[k]);
    const double me = ((u1[ns - 1] / u1[0]) + (u2[ns - 1] / u2[0])) / 2.;
    for (int fd = 0; fd < d; ++fd) {
        f[0 * d + fd] = mrho * mv[fd];
        for (int vd = 0; vd < d; ++vd) f[(1 + vd) * d + fd] = mrho * mv[fd] * mv[vd];
        f[(1 + fd) * d + fd] += mp;
        f[(ns - 1) * d + fd] = mrho * mv[fd] * me + mp * mv[fd];
    }
}

static void flux_ir(const dgo_ctx *c, const double *u1, const double *u2, double *f)
{
    /* euler.cpp:614-707 */
    const int d = c->dim, ns = c->ns;
    double z1[NS_MAX], z2[NS_MAX], avg[NS_MAX], lm[NS_MAX];
    {
        const double p1 = dgo_pressure(c, u1), p2 = dgo_pressure(c, u2);
        /* parameter vector from primitive (rho, vel, p) */
        z1[0] = sqrt(u1[0] / p1);
        z2[0] = sqrt(u2[0] / p2);
        for (int k = 0; k < d; ++k) { z1[1 + k] = z1[0] * (u1[1 + k] / u1[0]); z2[1 + k] = z2[0] * (u2[1 + k] / u2[0]); }
        z1[ns - 1] = sqrt(u1[0] * p1);
        z2[ns - 1] = sqrt(u2[0] * p2);
    }
    for (int s = 0; s < ns; ++s) avg[s] = 0.5 * (z1[s] + z2[s]);
    for (int s = 0; s < ns; ++s) lm[s] = dgo_log_mean(z1[s], z2[s]);
    double mv[3];
    const double mrho = avg[0] * lm[ns - 1];
    for (int k = 0; k < d; ++k) mv[k] = avg[1 + k] / avg[0];
    const double mp = avg[ns - 1] / avg[0];
    double mh = (c->gam + 1.0) * (lm[ns - 1] / lm[0]) + c->gamm1 * mp;
    mh /= 2.0 * c->gam;
    mh *= c->gam / (mrho * c->gamm1);
    double vsq = 0.0;
    for (int k = 0; k < d; ++k) vsq += mv[k] * mv[k];
    mh += 0.5 * vsq;
    for (int fd = 0; fd < d; ++fd) {
        f[0 * d + fd] = mrho * mv[fd];
        for (int vd = 0; vd < d; ++vd) f[(1 + vd) * d + fd] = mrho * mv[fd] * mv[vd];
        f[(1 + fd) * d + fd] += mp;
        f[(ns - 1) * d + fd] = mrho * mv[fd] * mh;
    }
}

static void flux_ch_ra(const dgo_ctx *c, int ranocha, const double *u1, const double *u2, double *f)
{
    /* Chandrashekar euler.cpp:710-756, Ranocha :759-807 */
    const int d = c->dim, ns = c->ns;
    const double rho_log = dgo_log_mean(u1[0], u2[0]);
    const double p1 = dgo_pressure(c, u1), p2 = dgo_pressure(c, u2);
    const double b1 = ranocha ? u1[0] / p1 : u1[0] / (2.0 * p1);
    const double b2 = ranocha ? u2[0] / p2 : u2[0] / (2.0 * p2);
    const double beta_log = dgo_log_mean(b1, b2);
    double v1[3], v2[3], va[3];
    velocities(c, u1, v1);
    velocities(c, u2, v2);
    const double p_hat = ranocha ? 0.5 * (p1 + p2) : 0.5 * (u1[0] + u2[0]) / (2.0 * 0.5 * (b1 + b2));
    double vsq = 0.0;
    for (int k = 0; k < d; ++k) {
        va[k] = 0.5 * (v1[k] + v2[k]);
        vsq += (0.5 * (v1[k] + v2[k])) * (0.5 * (v1[k] + v2[k]));
    }
    double H = ranocha ? 1.0 / (beta_log * c->gamm1) + vsq + 2.0 * p_hat / rho_log
                       : 1.0 / (2.0 * beta_log * c->gamm1) + vsq + p_hat / rho_log;
    for (int k = 0; k < d; ++k) H -= 0.5 * (0.5 * (v1[k] * v1[k] + v2[k] * v2[k]));
    for (int fd = 0; fd < d; ++fd) {
        f[0 * d + fd] = rho_log * va[fd];
        for (int vd = 0; vd < d; ++vd) f[(1 + vd) * d + fd] = rho_log * va[fd] * va[vd];
        f[(1 + fd) * d + fd] += p_hat;
        f[(ns - 1) * d + fd] = rho_log * va[fd] * H;
        if (ranocha) f[(ns - 1) * d + fd] -= (0.5 * (p1 * v1[fd] + p2 * v2[fd]));
    }
}

void dgo_two_point_flux(const dgo_ctx *c, int which, const double *u1, const double *u2, double *f)
{
    /* convective_numerical_split_flux dispatch (euler.cpp:567-584) */
    switch (which) {
    case DGO_KG: flux_kg(c, u1, u2, f); break;
    case DGO_IR: flux_ir(c, u1, u2, f); break;
    case DGO_CH: flux_ch_ra(c, 0, u1, u2, f); break;
    default: flux_ch_ra(c, 1, u1, u2, f); break;
    }
}

double dgo_max_convective_eigenvalue(const dgo_ctx *c, const double *u)
{
    /* euler.cpp:1033-1046 */
    double vel[3], vel2 = 0.0;
    velocities(c, u, vel);
    const double sound = sqrt(dgo_pressure(c, u) * c->gam / u[0]);
    for (int k = 0; k < c->dim; ++k) vel2 = vel2 + vel[k] * vel[k];
    return sqrt(vel2) + sound;
}

static double max_normal_eig(const dgo_ctx *c, const double *u, const double *n)
{
    /* euler.cpp:1049-1062 */
    double vel[3];
    velocities(c, u, vel);
    const double sound = sqrt(dgo_pressure(c, u) * c->gam / u[0]);
    double vn = 0.0;
    for (int k = 0; k < c->dim; ++k) vn += vel[k] * n[k];
    return fabs(vn) + sound;
}

static void roe_dissipation(const dgo_ctx *c, const double *ui, const double *ue, const double *n, double *out)
{
    /* RoeBaseRiemannSolverDissipation (convective_numerical_flux.cpp:332-480) with Harten fix (:216-232) */
    const int d = c->dim, ns = c->ns;
    const double rL = ui[0], rR = ue[0];
    double vL[3], vR[3];
    velocities(c, ui, vL);
    velocities(c, ue, vR);
    const double pL = dgo_pressure(c, ui), pR = dgo_pressure(c, ue);
    double vnL = 0.0, vnR = 0.0;
    for (int k = 0; k < d; ++k) vnL += vL[k] * n[k];
    for (int k = 0; k < d; ++k) vnR += vR[k] * n[k];
    const double hL = (ui[ns - 1] + pL) / rL;
    const double hR = (ue[ns - 1] + pR) / rR;
    const double r = sqrt(rR / rL);
    const double rp1 = r + 1.0;
    const double rho_ravg = r * rL;
    double v_ravg[3];
    for (int k = 0; k < d; ++k) v_ravg[k] = (r * vR[k] + vL[k]) / rp1;
    const double h_ravg = (r * hR + hL) / rp1;
    double vel2_ravg = 0.0;
    for (int k = 0; k < d; ++k) vel2_ravg = vel2_ravg + v_ravg[k] * v_ravg[k];
    double vn_ravg = 0.0;
    for (int k = 0; k < d; ++k) vn_ravg += v_ravg[k] * n[k];
    const double c2 = c->gamm1 * (h_ravg - 0.5 * vel2_ravg);
    double c_ravg = 1e10;
    if (c2 > 0.0) c_ravg = sqrt(c2);
    double eig[3] = {fabs(vn_ravg - c_ravg), fabs(vn_ravg), fabs(vn_ravg + c_ravg)};
    const double cL = sqrt(pL * c->gam / rL), cR = sqrt(pR * c->gam / rR);
    const double eL[3] = {fabs(vnL - cL), fabs(vnL), fabs(vnL + cL)};
    const double eR[3] = {fabs(vnR - cR), fabs(vnR), fabs(vnR + cR)};
    const double dp = pR - pL, drho = rR - rL;
    const double dVn = vnR - vnL;
    double dVt[3];
    for (int k = 0; k < d; ++k) dVt[k] = (vR[k] - vL[k]) - dVn * n[k];
    for (int e = 0; e < 3 && g_mutation != DGO_MUT_ROE_NO_HARTEN; ++e) {
        const double a1 = fabs(eig[e] - eL[e]), a2 = fabs(eR[e] - eig[e]);
        const double eps = a1 > a2 ? a1 : a2;
        if (eig[e] < eps) eig[e] = 0.5 * (eig[e] * eig[e] / eps + eps);
    }
    double co[4];
    co[0] = eig[0] * (dp - rho_ravg * c_ravg * dVn) / (2.0 * c2);
    co[1] = eig[1] * (drho - dp / c2);
    co[2] = (g_mutation == DGO_MUT_ROE_NO_SHEAR) ? 0.0 : eig[1] * rho_ravg;
    co[3] = eig[2] * (dp + rho_ravg * c_ravg * dVn) / (2.0 * c2);
    double A[NS_MAX];
    A[0] = co[0] * 1.0;
    for (int k = 0; k < d; ++k) A[1 + k] = co[0] * (v_ravg[k] - c_ravg * n[k]);
    A[ns - 1] = co[0] * (h_ravg - c_ravg * vn_ravg);
    A[0] += co[1] * 1.0;
    for (int k = 0; k < d; ++k) A[1 + k] += co[1] * v_ravg[k];
    A[ns - 1] += co[1] * vel2_ravg * 0.5;
    A[0] += co[2] * 0.0;
    double dvt_dot = 0.0;
    for (int k = 0; k < d; ++k) { A[1 + k] += co[2] * dVt[k]; dvt_dot += v_ravg[k] * dVt[k]; }
    A[ns - 1] += co[2] * dvt_dot;
    A[0] += co[3] * 1.0;
    for (int k = 0; k < d; ++k) A[1 + k] += co[3] * (v_ravg[k] + c_ravg * n[k]);
    A[ns - 1] += co[3] * (h_ravg + c_ravg * vn_ravg);
    const double half = (g_mutation == DGO_MUT_ROE_NO_HALF) ? 1.0 : 0.5;
    for (int s = 0; s < ns; ++s) out[s] = -half * A[s];
}

void dgo_numerical_flux(const dgo_ctx *c, int which, const double *ui, const double *ue,
                        const double *n, double *out)
{
    /* NumericalFluxConvective::evaluate_flux = baseline + dissipation (convective_numerical_flux.cpp:34-54) */
    const int d = c->dim, ns = c->ns;
    double base[NS_MAX], diss[NS_MAX];
    double f1[NS_MAX * 3], f2[NS_MAX * 3];
    if (which == DGO_TWO_POINT || which == DGO_TWO_POINT_LXF || which == DGO_TWO_POINT_ROE) {
        dgo_two_point_flux(c, c->P.two_point_flux, ui, ue, f1);
        for (int s = 0; s < ns; ++s) {
            double fd = 0.0;
            for (int k = 0; k < d; ++k) fd += f1[s * d + k] * n[k];
            base[s] = fd;
        }
    } else {
        dgo_convective_flux(c, ui, f1);
        dgo_convective_flux(c, ue, f2);
        for (int s = 0; s < ns; ++s) {
            double fd = 0.0;
            for (int k = 0; k < d; ++k) fd += (0.5 * (f1[s * d + k] + f2[s * d + k])) * n[k];
            base[s] = fd;
        }
    }
    if (which == DGO_LXF || which == DGO_TWO_POINT_LXF) {
        /* LaxFriedrichsRiemannSolverDissipation (convective_numerical_flux.cpp:188-213) */
        const double ei = max_normal_eig(c, ui, n), ee = max_normal_eig(c, ue, n);
        double lam = (ei > ee) ? ei : ee;
        if (g_mutation == DGO_MUT_LXF_MIN) lam = (ei > ee) ? ee : ei;
        const double half = (g_mutation == DGO_MUT_LXF_NO_HALF) ? 1.0 : 0.5;
        for (int s = 0; s < ns; ++s) diss[s] = -half * lam * (ue[s] - ui[s]);
    } else if (which == DGO_ROE || which == DGO_TWO_POINT_ROE) {
        roe_dissipation(c, ui, ue, n, diss);
    } else {
        for (int s = 0; s < ns; ++s) diss[s] = 0.0;
    }
    for (int s = 0; s < ns; ++s) out[s] = base[s] + diss[s];
}

/* ------------------------------------------------------------------------ */
/* Navier-Stokes viscous flux (src/physics/navier_stokes.cpp)               */
/* ------------------------------------------------------------------------ */
static void diss_flux_impl(const dgo_ctx *c, const double *u, const double *g, int adiabatic, double *f);

void dgo_dissipative_flux(const dgo_ctx *c, const double *u, const double *g, double *f)
{
    diss_flux_impl(c, u, g, 0, f);
}

/* adiabatic: the heat flux set to zero (dissipative_flux_dot_normal_on_adiabatic_boundary,
   navier_stokes.cpp:1266-1310: a no-slip wall with the adiabatic thermal condition) */
static void diss_flux_impl(const dgo_ctx *c, const double *u, const double *g, int adiabatic, double *f)
{
    /* dissipative_flux_templated (navier_stokes.cpp:1197-1220) ; g[s*dim+k], f[s*dim+k] */
    const int d = c->dim, ns = c->ns;
    /* Step 1: primitive (euler.cpp:177-194) */
    const double rho = u[0];
    double vel[3];
    velocities(c, u, vel);
    const double p = dgo_pressure(c, u);
    double prim[NS_MAX];
    prim[0] = rho;
    for (int k = 0; k < d; ++k) prim[1 + k] = vel[k];
    prim[ns - 1] = p;
    /* Step 2: primitive gradient (euler.cpp:258-303, formulation 2) */
    double pg[NS_MAX][3];
    for (int k = 0; k < d; ++k) pg[0][k] = g[0 * d + k];
    for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) pg[1 + a][b] = (g[(1 + a) * d + b] - vel[a] * g[0 * d + b]) / rho;
    for (int a = 0; a < d; ++a) {
        pg[ns - 1][a] = g[(ns - 1) * d + a];
        for (int b = 0; b < d; ++b)
            pg[ns - 1][a] -= 0.5 * (prim[1 + b] * g[(1 + b) * d + a] + u[1 + b] * pg[1 + b][a]);
        pg[ns - 1][a] *= c->gamm1;
    }
    /* Step 3: stress (navier_stokes.cpp:754-830) */
    double S[3][3];
    for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) S[a][b] = 0.5 * (pg[1 + a][b] + pg[1 + b][a]);
    const double T = c->gam * c->mach2 * (p / rho); /* euler.cpp compute_temperature */
    double mu;
    if (c->P.use_constant_viscosity) mu = 1.0 * c->P.constant_viscosity;
    else mu = ((1.0 + c->temperature_ratio) / (T + c->temperature_ratio)) * pow(T, 1.5);
    const double mu_hat = mu / c->P.reynolds_inf;
    double divv = 0.0;
    for (int k = 0; k < d; ++k) divv += S[k][k];
    double tau[3][3];
    const double mu2 = (-2.0 / 3.0) * mu_hat;
    for (int a = 0; a < d; ++a) {
        for (int b = 0; b < d; ++b) tau[a][b] = 2.0 * mu_hat * S[a][b];
        tau[a][a] += mu2 * divv;
    }
    /* heat flux (navier_stokes.cpp:336-385) */
    const double kappa = mu_hat / (c->gamm1 * c->mach2 * c->P.prandtl);
    double q[3];
    for (int k = 0; k < d; ++k) {
        const double dT = (c->gam * c->mach2 * pg[ns - 1][k] - T * pg[0][k]) / rho;
        q[k] = adiabatic ? 0.0 : -kappa * dT;
    }
    /* Step 4 (navier_stokes.cpp:1321-1350) */
    for (int fd = 0; fd < d; ++fd) {
        f[0 * d + fd] = 0.0;
        for (int sd = 0; sd < d; ++sd) f[(1 + sd) * d + fd] = -tau[sd][fd];
        f[(ns - 1) * d + fd] = 0.0;
        for (int sd = 0; sd < d; ++sd) f[(ns - 1) * d + fd] -= vel[sd] * tau[fd][sd];
        f[(ns - 1) * d + fd] += q[fd];
    }
}

static void diss_flux_dot_n(const dgo_ctx *c, const double *u, const double *g, const double *n, double *out,
                            int adiabatic)
{
    double f[NS_MAX * 3];
    diss_flux_impl(c, u, g, adiabatic, f);
    for (int s = 0; s < c->ns; ++s) {
        out[s] = 0.0;
        for (int k = 0; k < c->dim; ++k) out[s] += f[s * c->dim + k] * n[k];
    }
}

/* SIPG / BR2 evaluate_auxiliary_flux (viscous_numerical_flux.cpp:144-328); adiabatic: on an
   adiabatic no-slip wall every dissipative flux of the formula drops its heat flux */
static void viscous_num_flux(const dgo_ctx *c, const double *ui, const double *ue, const double *gi,
                             const double *ge, const double *n, double penalty, double *out, int adiabatic)
{
    const int ns = c->ns, d = c->dim;
    double a[NS_MAX], b[NS_MAX];
    diss_flux_dot_n(c, ui, gi, n, a, adiabatic);
    diss_flux_dot_n(c, ue, ge, n, b, adiabatic);
    double avg[NS_MAX];
    for (int s = 0; s < ns; ++s) avg[s] = 0.5 * (a[s] + b[s]);
    if (c->P.diss_flux == DGO_SIPG) {
        double jump[NS_MAX * 3];
        for (int s = 0; s < ns; ++s)
            for (int k = 0; k < d; ++k) jump[s * d + k] = (ui[s] - ue[s]) * n[k];
        double ja[NS_MAX], jb[NS_MAX];
        diss_flux_dot_n(c, ui, jump, n, ja, adiabatic);
        diss_flux_dot_n(c, ue, jump, n, jb, adiabatic);
        for (int s = 0; s < ns; ++s) out[s] = avg[s] - penalty * (0.5 * (ja[s] + jb[s]));
    } else {
        for (int s = 0; s < ns; ++s) out[s] = avg[s];
    }
}

/* ------------------------------------------------------------------------ */
/* residual assembly                                                        */
/* ------------------------------------------------------------------------ */
static inline void metric_cofactor(const dgo_ctx *c, double C[3][3])
{
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) C[a][b] = (a == b && a < c->dim) ? c->cof : 0.0;
}

/* the metric cofactor at node q of a cell (constant on the affine box) */
static inline void metric_at(const dgo_ctx *c, int cell, int q, double C[3][3])
{
    if (!c->curv) {
        metric_cofactor(c, C);
        return;
    }
    const double *m = c->metC + ((long)cell * c->Np + q) * 9;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) C[a][b] = m[a * 3 + b];
}

static inline double jac_at(const dgo_ctx *c, int cell, int q)
{
    return c->curv ? c->metJ[(long)cell * c->Np + q] : c->detJ;
}

/* volume node of face node f of a face (GLL: the face nodes are volume nodes) */
static int face_vol_node(const dgo_ctx *c, int face, int f)
{
    const int n = c->n, dir = face / 2, end = (face % 2) ? n - 1 : 0;
    const int a = f % n, b = (c->dim == 3) ? f / n : 0;
    if (c->dim == 2) return dir == 0 ? end + n * a : a + n * end;
    if (dir == 0) return end + n * (a + n * b);
    if (dir == 1) return a + n * (end + n * b);
    return a + n * (b + n * end);
}

/* transform_physical_to_reference (operators.cpp:2354-2365): ref_i = sum_j C_ji phys_j */
static inline double phys_to_ref(const dgo_ctx *c, const double C[3][3], const double *phys, int i)
{
    double r = 0.0;
    for (int j = 0; j < c->dim; ++j) r += C[j][i] * phys[j];
    return r;
}

static void column_index(const dgo_ctx *c, int q, int l, int rd, int *col)
{
    *col = (int)c->had_cols[((unsigned)q * c->nq + l) * c->dim + rd];
}

/* volume term (strong_dg.cpp:1017-1579) */
static void volume_term(dgo_ctx *c, int cell, const double *u, double *lrhs)
{
    const int ns = c->ns, Np = c->Np, nq = c->nq, n = c->n, d = c->dim;
    const int split = c->P.use_split_form;
    const int ns_ = ns;
    double *coef = (double *)malloc(sizeof(double) * ns * Np);
    double *sq = (double *)malloc(sizeof(double) * ns * Np);
    double *aq = (double *)calloc((size_t)ns * d * Np, sizeof(double));
    double *pev = split ? (double *)malloc(sizeof(double) * ns * Np) : NULL;
    double *F2 = split ? (double *)calloc((size_t)ns * d * Np * nq, sizeof(double)) : NULL;
    double *Fc = (double *)calloc((size_t)ns * d * Np, sizeof(double));
    double *Fd = (double *)calloc((size_t)ns * d * Np, sizeof(double));
    double *tmp = (double *)malloc(sizeof(double) * Np * nq);
    double *tmp2 = (double *)malloc(sizeof(double) * Np);
    for (int s = 0; s < ns; ++s) {
        for (int i = 0; i < Np; ++i) coef[s * Np + i] = u[IDX(c, cell, s, i)];
        vol_apply(c, c->chi, nq, n, coef + s * Np, sq + s * Np, 0, 1.0);
        if (c->P.pde == DGO_NAVIER_STOKES)
            for (int k = 0; k < d; ++k) {
                for (int i = 0; i < Np; ++i) tmp2[i] = c->aux[k][IDX(c, cell, s, i)];
                vol_apply(c, c->chi, nq, n, tmp2, aq + (s * d + k) * Np, 0, 1.0);
            }
    }
    if (split) {
        /* entropy projection (strong_dg.cpp:1238-1263) */
        double *ev = (double *)malloc(sizeof(double) * ns * Np);
        for (int q = 0; q < Np; ++q) {
            double us[NS_MAX], vs[NS_MAX];
            for (int s = 0; s < ns; ++s) us[s] = sq[s * Np + q];
            dgo_entropy_variables(c, us, vs);
            for (int s = 0; s < ns; ++s) ev[s * Np + q] = vs[s];
        }
        for (int s = 0; s < ns; ++s) {
            vol_apply(c, c->proj, n, nq, ev + s * Np, tmp2, 0, 1.0);
            vol_apply(c, c->chi, nq, n, tmp2, pev + s * Np, 0, 1.0);
        }
        free(ev);
    }
    for (int q = 0; q < Np; ++q) {
        double st[NS_MAX], ag[NS_MAX * 3];
        double C[3][3];
        metric_at(c, cell, q, C);
        for (int s = 0; s < ns; ++s) {
            st[s] = sq[s * Np + q];
            for (int k = 0; k < d; ++k) ag[s * d + k] = aq[(s * d + k) * Np + q];
        }
        if (split) {
            double evq[NS_MAX];
            for (int s = 0; s < ns; ++s) evq[s] = pev[s * Np + q];
            dgo_conservative_from_entropy(c, evq, st);
            for (int l = 0; l < nq; ++l) {
                for (int rd = 0; rd < d; ++rd) {
                    int col;
                    column_index(c, q, l, rd, &col);
                    double evc[NS_MAX], uc[NS_MAX], F[NS_MAX * 3];
                    for (int s = 0; s < ns; ++s) evc[s] = pev[s * Np + col];
                    dgo_conservative_from_entropy(c, evc, uc);
                    dgo_two_point_flux(c, c->P.two_point_flux, st, uc, F);
                    /* the two-point average of the metric cofactor (strong_dg.cpp:1348-1372) */
                    double Cc[3][3], Cs[3][3];
                    metric_at(c, cell, g_mutation == DGO_MUT_NO_METRIC_AVERAGE ? q : col, Cc);
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) Cs[a][b] = 0.5 * (C[a][b] + Cc[a][b]);
                    for (int s = 0; s < ns; ++s)
                        F2[((s * d) + rd) * Np * nq + q * nq + l] = phys_to_ref(c, Cs, F + s * d, rd);
                }
            }
        } else {
            double F[NS_MAX * 3];
            dgo_convective_flux(c, st, F);
            for (int s = 0; s < ns; ++s)
                for (int k = 0; k < d; ++k) Fc[(s * d + k) * Np + q] = phys_to_ref(c, C, F + s * d, k);
        }
        if (c->P.pde == DGO_NAVIER_STOKES) {
            double Fv[NS_MAX * 3];
            dgo_dissipative_flux(c, st, ag, Fv);
            for (int s = 0; s < ns; ++s)
                for (int k = 0; k < d; ++k) Fd[(s * d + k) * Np + q] = phys_to_ref(c, C, Fv + s * d, k);
        }
    }
    /* sum-factorized Hadamard basis (operators.cpp:1014-1065) */
    double *Ssp = NULL;
    if (split) {
        Ssp = (double *)malloc(sizeof(double) * d * Np * nq);
        const int total = Np * nq;
        for (int idx = 0, counter = 0; idx < total; ++idx, ++counter) {
            if (counter == nq) counter = 0;
            const unsigned *r = c->had_rows + (size_t)idx * d;
            const unsigned *cc = c->had_cols + (size_t)idx * d;
            const unsigned N1 = (unsigned)nq;
            if (d == 2) {
                Ssp[(0 * Np + r[0]) * nq + counter] = c->skew[(r[0] % N1) * nq + (cc[0] % N1)] * c->wq[r[1] / N1];
                Ssp[(1 * Np + r[1]) * nq + counter] = c->skew[(r[1] / N1) * nq + (cc[1] / N1)] * c->wq[r[0] % N1];
            } else {
                Ssp[(0 * Np + r[0]) * nq + counter] = c->skew[(r[0] % N1) * nq + (cc[0] % N1)]
                                                      * c->wq[(r[1] / N1) % N1] * c->wq[r[2] / N1 / N1];
                Ssp[(1 * Np + r[1]) * nq + counter] = c->skew[((r[1] / N1) % N1) * nq + ((cc[1] / N1) % N1)]
                                                      * c->wq[r[0] % N1] * c->wq[r[2] / N1 / N1];
                Ssp[(2 * Np + r[2]) * nq + counter] = c->skew[(r[2] / N1 / N1) * nq + (cc[2] / N1 / N1)]
                                                      * c->wq[r[0] % N1] * c->wq[(r[1] / N1) % N1];
            }
        }
    }
    double *div = (double *)malloc(sizeof(double) * Np);
    double *ddiv = (double *)malloc(sizeof(double) * Np);
    double *rhs = (double *)malloc(sizeof(double) * Np);
    for (int s = 0; s < ns_; ++s) {
        if (split) {
            for (int rd = 0; rd < d; ++rd) {
                for (int i = 0; i < Np * nq; ++i)
                    tmp[i] = Ssp[(rd * Np) * nq + i] * F2[((s * d) + rd) * Np * nq + i];
                for (int q = 0; q < Np; ++q) {
                    if (rd == 0) div[q] = 0.0;
                    for (int l = 0; l < nq; ++l) div[q] += tmp[q * nq + l];
                }
            }
        } else {
            for (int k = 0; k < d; ++k)
                dir_apply(c, k, c->fb_grad, c->fb_vol, nq, nq, Fc + (s * d + k) * Np, div, k > 0, 1.0);
        }
        for (int k = 0; k < d; ++k)
            dir_apply(c, k, c->fb_grad, c->fb_vol, nq, nq, Fd + (s * d + k) * Np, ddiv, k > 0, 1.0);
        vol_inner(c, div, split ? c->ones_vol : c->wvol, rhs, 0, -1.0);
        vol_inner(c, ddiv, c->wvol, rhs, 1, -1.0);
        for (int i = 0; i < Np; ++i) lrhs[s * Np + i] += rhs[i];
    }
    free(coef); free(sq); free(aq); free(pev); free(F2); free(Fc); free(Fd); free(tmp); free(tmp2);
    free(Ssp); free(div); free(ddiv); free(rhs);
}

/* cell->extent_in_direction(k): the longest of the cell's edges along reference direction k, from
   its vertices (the corner nodes); h on the affine box */
static double cell_extent(const dgo_ctx *c, int cell, int k)
{
    if (!c->curv) return c->h;
    const int n = c->n, d = c->dim;
    double ext = 0.0;
    for (int a = 0; a < (1 << (d - 1)); ++a) {
        int lo[3] = {0, 0, 0}, bit = 0;
        for (int m = 0; m < d; ++m) {
            if (m == k) continue;
            lo[m] = ((a >> bit) & 1) ? n - 1 : 0;
            ++bit;
        }
        int hi[3] = {lo[0], lo[1], lo[2]};
        hi[k] = n - 1;
        const int q0 = lo[0] + n * (lo[1] + n * lo[2]), q1 = hi[0] + n * (hi[1] + n * hi[2]);
        const double *x0 = c->xn + ((long)cell * c->Np + q0) * 3, *x1 = c->xn + ((long)cell * c->Np + q1) * 3;
        double l2 = 0.0;
        for (int m = 0; m < 3; ++m) l2 += (x1[m] - x0[m]) * (x1[m] - x0[m]);
        const double l = sqrt(l2);
        if (l > ext) ext = l;
    }
    return ext;
}

/* evaluate_penalty_scaling (dg_base.cpp:397-413) of a cell's faces normal to direction k */
static double penalty_scaling(const dgo_ctx *c, int cell, int k)
{
    const unsigned deg = (unsigned)c->p;
    const unsigned degsq = (deg == 0) ? 1 : deg * (deg + 1);
    return degsq / cell_extent(c, cell, k) * c->P.sipg_penalty_factor;
}

/* face term (strong_dg.cpp:2368-3576); rhs_ext is the global vector */
/* Interior face (ce >= 0) or boundary face of type -ce (assemble_boundary_term_strong,
   strong_dg.cpp:1583-2364: the interior side's terms only, with the exterior state of the
   numerical flux from boundary_face_values of the interpolated interior trace). */
static void face_term(dgo_ctx *c, int ci, int fi, int ce, int fe, const double *u, double *lrhs, double *rhs_glob)
{
    const int ns = c->ns, Np = c->Np, nq = c->nq, n = c->n, d = c->dim, nfp = c->nfp;
    const int split = c->P.use_split_form;
    const int ns_visc = (c->P.pde == DGO_NAVIER_STOKES);
    /* a boundary face: the cell's own penalty (dg_base.cpp:563); an interior face: the average of both
       cells' (dg_base.cpp:615-617) */
    const double penalty = (ce < 0) ? penalty_scaling(c, ci, fi / 2)
                                    : 0.5 * (penalty_scaling(c, ci, fi / 2) + penalty_scaling(c, ce, fe / 2));
    const int boundary = ce < 0, nsides = boundary ? 1 : 2;
    const int cells[2] = {ci, ce}, faces[2] = {fi, fe};
    double *svol[2] = {NULL, NULL}, *ssurf[2] = {NULL, NULL}, *avol[2] = {NULL, NULL}, *asurf[2] = {NULL, NULL};
    double *Fc_f[2] = {NULL, NULL}, *Fd_f[2] = {NULL, NULL};
    double *pev_vol[2] = {NULL, NULL}, *pev_surf[2] = {NULL, NULL};
    double *tmp = (double *)malloc(sizeof(double) * Np);
    double *tmp2 = (double *)malloc(sizeof(double) * Np);
    for (int side = 0; side < nsides; ++side) {
        const int cell = cells[side], face = faces[side];
        svol[side] = (double *)malloc(sizeof(double) * ns * Np);
        ssurf[side] = (double *)malloc(sizeof(double) * ns * nfp);
        avol[side] = (double *)calloc((size_t)ns * d * Np, sizeof(double));
        asurf[side] = (double *)calloc((size_t)ns * d * nfp, sizeof(double));
        Fc_f[side] = (double *)calloc((size_t)ns * nfp, sizeof(double));
        Fd_f[side] = (double *)calloc((size_t)ns * nfp, sizeof(double));
        for (int s = 0; s < ns; ++s) {
            for (int i = 0; i < Np; ++i) tmp[i] = u[IDX(c, cell, s, i)];
            vol_apply(c, c->chi, nq, n, tmp, svol[side] + s * Np, 0, 1.0);
            surf_apply(c, face, c->chi_surf[face % 2], c->chi, nq, n, tmp, ssurf[side] + s * nfp, 0, 1.0);
            if (ns_visc)
                for (int k = 0; k < d; ++k) {
                    for (int i = 0; i < Np; ++i) tmp2[i] = c->aux[k][IDX(c, cell, s, i)];
                    vol_apply(c, c->chi, nq, n, tmp2, avol[side] + (s * d + k) * Np, 0, 1.0);
                    surf_apply(c, face, c->chi_surf[face % 2], c->chi, nq, n, tmp2, asurf[side] + (s * d + k) * nfp, 0, 1.0);
                }
        }
        /* volume reference fluxes at volume nodes, normal component interpolated to the face (:2840-3041) */
        const int dnz = face / 2;
        const double nref = (face % 2) ? 1.0 : -1.0;
        double *Fc = (double *)calloc((size_t)ns * Np, sizeof(double));
        double *Fd = (double *)calloc((size_t)ns * Np, sizeof(double));
        for (int q = 0; q < Np; ++q) {
            double st[NS_MAX], ag[NS_MAX * 3];
            double C[3][3];
            metric_at(c, cell, q, C);
            for (int s = 0; s < ns; ++s) {
                st[s] = svol[side][s * Np + q];
                for (int k = 0; k < d; ++k) ag[s * d + k] = avol[side][(s * d + k) * Np + q];
            }
            if (!split) {
                double F[NS_MAX * 3];
                dgo_convective_flux(c, st, F);
                for (int s = 0; s < ns; ++s) Fc[s * Np + q] = phys_to_ref(c, C, F + s * d, dnz);
            }
            if (ns_visc) {
                double Fv[NS_MAX * 3];
                dgo_dissipative_flux(c, st, ag, Fv);
                for (int s = 0; s < ns; ++s) Fd[s * Np + q] = phys_to_ref(c, C, Fv + s * d, dnz);
            }
        }
        for (int s = 0; s < ns; ++s) {
            if (!split)
                surf_apply(c, face, c->fb_surf[face % 2], c->fb_vol, nq, nq, Fc + s * Np, Fc_f[side] + s * nfp, 0, nref);
            surf_apply(c, face, c->fb_surf[face % 2], c->fb_vol, nq, nq, Fd + s * Np, Fd_f[side] + s * nfp, 0, nref);
        }
        free(Fc); free(Fd);
        if (split) {
            /* entropy variables and projection (:3050-3142) */
            pev_vol[side] = (double *)malloc(sizeof(double) * ns * Np);
            pev_surf[side] = (double *)malloc(sizeof(double) * ns * nfp);
            double *ev = (double *)malloc(sizeof(double) * ns * Np);
            for (int q = 0; q < Np; ++q) {
                double us[NS_MAX], vs[NS_MAX];
                for (int s = 0; s < ns; ++s) us[s] = svol[side][s * Np + q];
                dgo_entropy_variables(c, us, vs);
                for (int s = 0; s < ns; ++s) ev[s * Np + q] = vs[s];
            }
            for (int s = 0; s < ns; ++s) {
                vol_apply(c, c->proj, n, nq, ev + s * Np, tmp2, 0, 1.0);
                vol_apply(c, c->chi, nq, n, tmp2, pev_vol[side] + s * Np, 0, 1.0);
                surf_apply(c, face, c->chi_surf[face % 2], c->chi, nq, n, tmp2, pev_surf[side] + s * nfp, 0, 1.0);
            }
            free(ev);
        }
    }
    /* surface-volume hybrid two-point terms (:3162-3350) */
    double *sv_surf[2] = {NULL, NULL}, *sv_vol[2] = {NULL, NULL};
    if (split) {
        for (int side = 0; side < nsides; ++side) {
            const int face = faces[side], dnz = face / 2;
            const double nref = (face % 2) ? 1.0 : -1.0;
            const unsigned *srow = c->had_srows[dnz], *scol = c->had_scols[dnz];
            double *S2 = (double *)malloc(sizeof(double) * ns * nfp * nq);
            for (int f = 0; f < nfp; ++f) {
                double evf[NS_MAX], uf[NS_MAX];
                for (int s = 0; s < ns; ++s) evf[s] = pev_surf[side][s * nfp + f];
                dgo_conservative_from_entropy(c, evf, uf);
                for (int l = 0; l < nq; ++l) {
                    const int row_index = f * nq + l;
                    if ((int)srow[row_index] != f) { fprintf(stderr, "surface Hadamard pattern mismatch\n"); abort(); }
                    const int iv = (int)scol[row_index];
                    double evv[NS_MAX], uv[NS_MAX], F[NS_MAX * 3];
                    for (int s = 0; s < ns; ++s) evv[s] = pev_vol[side][s * Np + iv];
                    dgo_conservative_from_entropy(c, evv, uv);
                    dgo_two_point_flux(c, c->P.two_point_flux, uv, uf, F);
                    /* metric average of the face node and the volume node (strong_dg.cpp:3172-3200) */
                    double Cf[3][3], Cv[3][3], Cs[3][3];
                    metric_at(c, cells[side], face_vol_node(c, face, f), Cf);
                    metric_at(c, cells[side], iv, Cv);
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) Cs[a][b] = 0.5 * (Cf[a][b] + Cv[a][b]);
                    for (int s = 0; s < ns; ++s) S2[(s * nfp + f) * nq + l] = phys_to_ref(c, Cs, F + s * d, dnz);
                }
            }
            /* sum_factorized_Hadamard_surface_basis_assembly (operators.cpp:1132-1196) */
            double *B = (double *)malloc(sizeof(double) * nfp * nq);
            const unsigned N1 = (unsigned)nq;
            for (int idx = 0, counter = 0; idx < nfp * nq; ++idx, ++counter) {
                if (counter == nq) counter = 0;
                const unsigned r = srow[idx], cc = scol[idx];
                double v;
                if (d == 2) {
                    if (dnz == 0) v = c->fb_surf[face % 2][cc % N1] * c->wq[r % N1];
                    else v = c->fb_surf[face % 2][cc / N1] * c->wq[r % N1];
                } else {
                    if (dnz == 0) v = c->fb_surf[face % 2][cc % N1] * c->wq[r % N1] * c->wq[r / N1];
                    else if (dnz == 1) v = c->fb_surf[face % 2][(cc / N1) % N1] * c->wq[r % N1] * c->wq[r / N1];
                    else v = c->fb_surf[face % 2][cc / N1 / N1] * c->wq[r % N1] * c->wq[r / N1];
                }
                B[r * nq + counter] = v;
            }
            sv_surf[side] = (double *)calloc((size_t)ns * nfp, sizeof(double));
            sv_vol[side] = (double *)calloc((size_t)ns * Np, sizeof(double));
            for (int s = 0; s < ns; ++s)
                for (int f = 0; f < nfp; ++f)
                    for (int l = 0; l < nq; ++l) {
                        const double H = B[f * nq + l] * S2[(s * nfp + f) * nq + l];
                        sv_surf[side][s * nfp + f] -= H * nref;
                        sv_vol[side][s * Np + scol[f * nq + l]] += H * nref;
                    }
            free(S2); free(B);
        }
    }
    /* numerical fluxes (:3354-3452) */
    double *cnf = (double *)malloc(sizeof(double) * ns * nfp);
    double *dnf = (double *)calloc((size_t)ns * nfp, sizeof(double));
    {
        const int dnz = fi / 2;
        double nref[3] = {0, 0, 0};
        nref[dnz] = (fi % 2) ? 1.0 : -1.0;
        for (int f = 0; f < nfp; ++f) {
            double ui[NS_MAX], ue[NS_MAX];
            double nphys[3];
            /* the interior cell's metric at the face node (metric_oper_int) */
            double C[3][3];
            metric_at(c, ci, face_vol_node(c, fi, f), C);
            for (int a = 0; a < d; ++a) {
                nphys[a] = 0.0;
                for (int b = 0; b < d; ++b) nphys[a] += C[a][b] * nref[b];
            }
            double nrm = 0.0;
            for (int a = 0; a < d; ++a) nrm += nphys[a] * nphys[a];
            nrm = sqrt(nrm);
            for (int a = 0; a < d; ++a) nphys[a] /= nrm;
            if (boundary) {
                /* the interpolated interior trace (use_projected_entropy_variables_for_nsfr_boundary_term
                   defaults to false) and its boundary state */
                for (int s = 0; s < ns; ++s) ui[s] = ssurf[0][s * nfp + f];
                dgo_boundary_state(c, -ce, ui, nphys, ue);
            } else if (split) {
                double vi[NS_MAX], ve[NS_MAX];
                for (int s = 0; s < ns; ++s) { vi[s] = pev_surf[0][s * nfp + f]; ve[s] = pev_surf[1][s * nfp + f]; }
                dgo_conservative_from_entropy(c, vi, ui);
                dgo_conservative_from_entropy(c, ve, ue);
            } else {
                for (int s = 0; s < ns; ++s) { ui[s] = ssurf[0][s * nfp + f]; ue[s] = ssurf[1][s * nfp + f]; }
            }
            double fl[NS_MAX];
            dgo_numerical_flux(c, c->P.conv_flux, ui, ue, nphys, fl);
            for (int s = 0; s < ns; ++s) cnf[s * nfp + f] = nrm * fl[s];
            if (ns_visc) {
                double si[NS_MAX], se[NS_MAX], gi[NS_MAX * 3], ge[NS_MAX * 3], vf[NS_MAX];
                for (int s = 0; s < ns; ++s) {
                    si[s] = ssurf[0][s * nfp + f];
                    for (int k = 0; k < d; ++k) gi[s * d + k] = asurf[0][(s * d + k) * nfp + f];
                    if (boundary) continue;
                    se[s] = ssurf[1][s * nfp + f];
                    for (int k = 0; k < d; ++k) ge[s * d + k] = asurf[1][(s * d + k) * nfp + f];
                }
                /* a wall: the viscous boundary state and gradient (strong_dg.cpp:2276-2300) */
                if (boundary) visc_boundary_state(c, -ce, si, gi, se, ge);
                viscous_num_flux(c, si, se, gi, ge, nphys, penalty, vf, boundary && wall_adiabatic(c, -ce));
                for (int s = 0; s < ns; ++s) dnf[s * nfp + f] = nrm * vf[s];
            }
        }
    }
    /* lift into interior and exterior (:3454-3575) */
    double *r = (double *)malloc(sizeof(double) * Np);
    for (int side = 0; side < nsides; ++side) {
        const int face = faces[side];
        const double sg = side == 0 ? -1.0 : 1.0;
        for (int s = 0; s < ns; ++s) {
            if (split) {
                surf_inner(c, face, sv_surf[side] + s * nfp, c->ones_face, r, c->chi_surf[face % 2], c->chi, nq, n, 0, -1.0);
                vol_inner(c, sv_vol[side] + s * Np, c->ones_vol, r, 1, -1.0);
            } else {
                surf_inner(c, face, Fc_f[side] + s * nfp, c->wface, r, c->chi_surf[face % 2], c->chi, nq, n, 0, 1.0);
            }
            surf_inner(c, face, Fd_f[side] + s * nfp, c->wface, r, c->chi_surf[face % 2], c->chi, nq, n, 1, 1.0);
            surf_inner(c, face, cnf + s * nfp, c->wface, r, c->chi_surf[face % 2], c->chi, nq, n, 1, sg);
            surf_inner(c, face, dnf + s * nfp, c->wface, r, c->chi_surf[face % 2], c->chi, nq, n, 1, sg);
            if (side == 0) {
                for (int i = 0; i < Np; ++i) lrhs[s * Np + i] += r[i];
            } else {
                for (int i = 0; i < Np; ++i) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                    rhs_glob[IDX(c, ce, s, i)] += r[i];
                }
            }
        }
    }
    free(r); free(cnf); free(dnf); free(tmp); free(tmp2);
    for (int side = 0; side < 2; ++side) {
        free(svol[side]); free(ssurf[side]); free(avol[side]); free(asurf[side]);
        free(Fc_f[side]); free(Fd_f[side]); free(pev_vol[side]); free(pev_surf[side]);
        free(sv_surf[side]); free(sv_vol[side]);
    }
}

/* auxiliary equation (strong_dg.cpp:498-1008) */
static void aux_volume(dgo_ctx *c, int cell, const double *u, double *laux /*[d][ns*Np]*/)
{
    const int ns = c->ns, Np = c->Np, nq = c->nq, n = c->n, d = c->dim;
    double *coef = (double *)malloc(sizeof(double) * Np);
    double *sq = (double *)malloc(sizeof(double) * Np);
    double *rg = (double *)malloc(sizeof(double) * d * Np);
    double *pg = (double *)malloc(sizeof(double) * Np);
    double *rhs = (double *)malloc(sizeof(double) * Np);
    for (int s = 0; s < ns; ++s) {
        for (int i = 0; i < Np; ++i) coef[i] = u[IDX(c, cell, s, i)];
        vol_apply(c, c->chi, nq, n, coef, sq, 0, 1.0);
        for (int k = 0; k < d; ++k) dir_apply(c, k, c->fb_grad, c->fb_vol, nq, nq, sq, rg + k * Np, 0, 1.0);
        for (int k = 0; k < d; ++k) {
            for (int q = 0; q < Np; ++q) {
                /* the physical gradient C grad_ref u (strong_dg.cpp:700-712) */
                double C[3][3];
                metric_at(c, cell, q, C);
                pg[q] = 0.0;
                for (int j = 0; j < d; ++j) pg[q] += C[k][j] * rg[j * Np + q];
            }
            vol_inner(c, pg, c->wvol, rhs, 0, 1.0);
            for (int i = 0; i < Np; ++i) laux[k * ns * Np + s * Np + i] += rhs[i];
        }
    }
    free(coef); free(sq); free(rg); free(pg); free(rhs);
}

/* ce < 0: boundary face of type -ce (assemble_boundary_term_auxiliary_equation, strong_dg.cpp:
   790-870): the exterior state is the viscous wall state of the interior trace, only the interior
   side is lifted */
static void aux_face(dgo_ctx *c, int ci, int fi, int ce, int fe, const double *u, double *laux)
{
    const int ns = c->ns, Np = c->Np, nq = c->nq, n = c->n, d = c->dim, nfp = c->nfp;
    const int boundary = ce < 0;
    double *tmp = (double *)malloc(sizeof(double) * Np);
    double *si = (double *)malloc(sizeof(double) * ns * nfp);
    double *se = (double *)malloc(sizeof(double) * ns * nfp);
    for (int s = 0; s < ns; ++s) {
        for (int i = 0; i < Np; ++i) tmp[i] = u[IDX(c, ci, s, i)];
        surf_apply(c, fi, c->chi_surf[fi % 2], c->chi, nq, n, tmp, si + s * nfp, 0, 1.0);
        if (boundary) continue;
        for (int i = 0; i < Np; ++i) tmp[i] = u[IDX(c, ce, s, i)];
        surf_apply(c, fe, c->chi_surf[fe % 2], c->chi, nq, n, tmp, se + s * nfp, 0, 1.0);
    }
    if (boundary)
        for (int f = 0; f < nfp; ++f) {
            double ui[NS_MAX], ub[NS_MAX], g0[NS_MAX * 3] = {0}, gb[NS_MAX * 3];
            for (int s = 0; s < ns; ++s) ui[s] = si[s * nfp + f];
            visc_boundary_state(c, -ce, ui, g0, ub, gb);
            for (int s = 0; s < ns; ++s) se[s * nfp + f] = ub[s];
        }
    double nref[3] = {0, 0, 0};
    nref[fi / 2] = (fi % 2) ? 1.0 : -1.0;
    double *vi = (double *)malloc(sizeof(double) * ns * d * nfp);
    double *ve = (double *)malloc(sizeof(double) * ns * d * nfp);
    for (int f = 0; f < nfp; ++f) {
        /* the interior cell's metric at the face node */
        double C[3][3];
        metric_at(c, ci, face_vol_node(c, fi, f), C);
        double nphys[3];
        for (int a = 0; a < d; ++a) {
            nphys[a] = 0.0;
            for (int b = 0; b < d; ++b) nphys[a] += C[a][b] * nref[b];
        }
        double nrm = 0.0;
        for (int a = 0; a < d; ++a) nrm += nphys[a] * nphys[a];
        nrm = sqrt(nrm);
        for (int a = 0; a < d; ++a) nphys[a] /= nrm;
        for (int s = 0; s < ns; ++s) {
            const double avg = 0.5 * (si[s * nfp + f] + se[s * nfp + f]);
            for (int k = 0; k < d; ++k) {
                vi[(s * d + k) * nfp + f] = (avg - si[s * nfp + f]) * nphys[k] * nrm;
                ve[(s * d + k) * nfp + f] = (avg - se[s * nfp + f]) * (-nphys[k]) * nrm;
            }
        }
    }
    double *r = (double *)malloc(sizeof(double) * Np);
    for (int s = 0; s < ns; ++s)
        for (int k = 0; k < d; ++k) {
            surf_inner(c, fi, vi + (s * d + k) * nfp, c->wface, r, c->chi_surf[fi % 2], c->chi, nq, n, 0, 1.0);
            for (int i = 0; i < Np; ++i) laux[k * ns * Np + s * Np + i] += r[i];
            if (boundary) continue;
            surf_inner(c, fe, ve + (s * d + k) * nfp, c->wface, r, c->chi_surf[fe % 2], c->chi, nq, n, 0, 1.0);
            for (int i = 0; i < Np; ++i) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                c->aux_rhs[k][IDX(c, ce, s, i)] += r[i];
            }
        }
    free(r); free(vi); free(ve); free(tmp); free(si); free(se);
}

static void assemble_auxiliary(dgo_ctx *c, const double *u)
{
    const long nd = dgo_n_dofs(c);
    const int d = c->dim, ns = c->ns, Np = c->Np;
    for (int k = 0; k < d; ++k) memset(c->aux_rhs[k], 0, sizeof(double) * nd);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads)
#endif
    for (int e = 0; e < c->ncells; ++e) {
        double *laux = (double *)calloc((size_t)d * ns * Np, sizeof(double));
        aux_volume(c, e, u, laux);
        for (int f = 0; f < 2 * d; ++f) {
            const int nb = c->nbr[e * 2 * d + f];
            if (e < nb || nb < 0) aux_face(c, e, f, nb, f ^ 1, u, laux);
        }
        for (int k = 0; k < d; ++k)
            for (int s = 0; s < ns; ++s)
                for (int i = 0; i < Np; ++i) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                    c->aux_rhs[k][IDX(c, e, s, i)] += laux[k * ns * Np + s * Np + i];
                }
        free(laux);
    }
    for (int k = 0; k < d; ++k) dgo_apply_inverse_mass(c, c->aux_rhs[k], c->aux[k], 1);
}

void dgo_assemble_residual(dgo_ctx *c, const double *u, double *rhs)
{
    /* DGBase::assemble_residual (dg_base.cpp:2599-2914) */
    const long nd = dgo_n_dofs(c);
    const int d = c->dim, ns = c->ns, Np = c->Np;
    memset(rhs, 0, sizeof(double) * nd);
    if (c->P.pde == DGO_NAVIER_STOKES) assemble_auxiliary(c, u);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads)
#endif
    for (int e = 0; e < c->ncells; ++e) {
        double *lrhs = (double *)calloc((size_t)ns * Np, sizeof(double));
        volume_term(c, e, u, lrhs);
        for (int f = 0; f < 2 * d; ++f) {
            const int nb = c->nbr[e * 2 * d + f];
            if (e < nb || nb < 0) face_term(c, e, f, nb, f ^ 1, u, lrhs, rhs);
        }
        for (int i = 0; i < ns * Np; ++i) {
#ifdef _OPENMP
#pragma omp atomic
#endif
            rhs[(long)e * ns * Np + i] += lrhs[i];
        }
        free(lrhs);
    }
}

void dgo_set_fsal(dgo_ctx *c, int on) { c->fsal = on != 0; }

/* Rounding scale of the residual, per DoF (test infrastructure, collocated GLL nodes): the sum
   of the absolute values of the terms the residual at that DoF is summed from -- two-point
   volume fluxes times the skew operator (split form) or flux values times the derivative
   operator (conservative form and the viscous flux), and at face nodes the numerical fluxes,
   the one-sided terms of their dissipation (the LxF/Roe dissipation -1/2 |A| (u_R - u_L) is the
   sum of -1/2 |A| u_R and +1/2 |A| u_L; bounded here by 1/2 lambda (|u_L| + |u_R|), lambda the
   larger spectral radius, which bounds every eigenvalue of the Roe matrix), the one-sided terms
   of the SIPG penalty flux (F_v(u, (u_i - u_e) n) is linear in the jump) and the interior fluxes
   of the strong form -- each with the node's quadrature weight and metric factor W_i h^(d-1)
   (surface terms also 1/w_0).  A residual computed in another summation order, or from states
   that differ by a few ulps (the entropy-variable round trip), differs from the oracle's by a
   few units of FP64 rounding of this scale. */
void dgo_residual_magnitude(dgo_ctx *c, const double *u, double *mag)
{
    const long nd = dgo_n_dofs(c);
    const int d = c->dim, ns = c->ns, Np = c->Np, n = c->n;
    const int split = c->P.use_split_form, visc = (c->P.pde == DGO_NAVIER_STOKES);
    memset(mag, 0, sizeof(double) * nd);
    if (c->P.flux_nodes != DGO_GLL) return;
    if (visc) assemble_auxiliary(c, u);
    double D[MAXN][MAXN], Sk[MAXN][MAXN];
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            D[a][b] = c->fb_grad[a * n + b];
            Sk[a][b] = c->skew[a * n + b] / c->wq[a];
        }
    const double hd1 = (d == 3) ? c->h * c->h : c->h;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads)
#endif
    for (int e = 0; e < c->ncells; ++e) {
        for (int i = 0; i < Np; ++i) {
            int ci[3] = {i % n, (i / n) % n, (d == 3) ? i / (n * n) : 0};
            double ui[NS_MAX], gi[NS_MAX * 3], acc[NS_MAX];
            for (int s = 0; s < ns; ++s) {
                ui[s] = u[IDX(c, e, s, i)];
                acc[s] = 0.0;
                for (int k = 0; k < d; ++k) gi[s * d + k] = visc ? c->aux[k][IDX(c, e, s, i)] : 0.0;
            }
            for (int k = 0; k < d; ++k) {
                for (int b = 0; b < n; ++b) {
                    int cj[3] = {ci[0], ci[1], ci[2]};
                    cj[k] = b;
                    const int j = cj[0] + n * (cj[1] + n * cj[2]);
                    double uj[NS_MAX], gj[NS_MAX * 3], F[NS_MAX * 3];
                    for (int s = 0; s < ns; ++s) {
                        uj[s] = u[IDX(c, e, s, j)];
                        for (int kk = 0; kk < d; ++kk) gj[s * d + kk] = visc ? c->aux[kk][IDX(c, e, s, j)] : 0.0;
                    }
                    /* reference-direction component of a physical flux: F_k on the affine box
                       (h^(d-1) applied below), sum_m F_m Cbar[m][k] on a curvilinear grid */
                    double Ci[3][3], Cj[3][3];
                    metric_at(c, e, i, Ci);
                    metric_at(c, e, j, Cj);
                    if (split) {
                        if (b != ci[k]) {
                            dgo_two_point_flux(c, c->P.two_point_flux, ui, uj, F);
                            for (int s = 0; s < ns; ++s) {
                                double fk = F[s * d + k];
                                if (c->curv) {
                                    fk = 0.0;
                                    for (int m = 0; m < d; ++m) fk += F[s * d + m] * 0.5 * (Ci[m][k] + Cj[m][k]);
                                }
                                acc[s] += fabs(Sk[ci[k]][b] * fk);
                            }
                        }
                    } else {
                        dgo_convective_flux(c, uj, F);
                        for (int s = 0; s < ns; ++s) {
                            double fk = F[s * d + k];
                            if (c->curv) {
                                fk = 0.0;
                                for (int m = 0; m < d; ++m) fk += F[s * d + m] * Cj[m][k];
                            }
                            acc[s] += fabs(D[ci[k]][b] * fk);
                        }
                    }
                    if (visc) {
                        dgo_dissipative_flux(c, uj, gj, F);
                        for (int s = 0; s < ns; ++s) {
                            double fk = F[s * d + k];
                            if (c->curv) {
                                fk = 0.0;
                                for (int m = 0; m < d; ++m) fk += F[s * d + m] * Cj[m][k];
                            }
                            acc[s] += fabs(D[ci[k]][b] * fk);
                        }
                    }
                }
                if (ci[k] != 0 && ci[k] != n - 1) continue;
                /* face node: numerical fluxes and the interior fluxes of the strong form */
                const int side = (ci[k] == n - 1), face = 2 * k + side;
                const int nb = c->nbr[e * 2 * d + face];
                const double penalty = (nb < 0) ? penalty_scaling(c, e, k)
                                                : 0.5 * (penalty_scaling(c, e, k) + penalty_scaling(c, nb, k));
                int cj[3] = {ci[0], ci[1], ci[2]};
                cj[k] = side ? 0 : n - 1;
                const int j = cj[0] + n * (cj[1] + n * cj[2]);
                double ue[NS_MAX], ge[NS_MAX * 3], nrm[3] = {0, 0, 0}, fl[NS_MAX], F[NS_MAX * 3];
                nrm[k] = side ? 1.0 : -1.0;
                double area = 1.0;  /* |C^T n_ref| on a curvilinear grid */
                if (c->curv) {
                    double Ci[3][3], a2 = 0.0;
                    metric_at(c, e, i, Ci);
                    for (int m = 0; m < d; ++m) {
                        nrm[m] = Ci[m][k] * (side ? 1.0 : -1.0);
                        a2 += nrm[m] * nrm[m];
                    }
                    area = sqrt(a2);
                    for (int m = 0; m < d; ++m) nrm[m] /= area;
                }
                if (nb < 0) {
                    dgo_boundary_state(c, -nb, ui, nrm, ue);
                    for (int s = 0; s < ns * 3; ++s) ge[s] = 0.0;
                } else {
                    for (int s = 0; s < ns; ++s) {
                        ue[s] = u[IDX(c, nb, s, j)];
                        for (int kk = 0; kk < d; ++kk) ge[s * d + kk] = visc ? c->aux[kk][IDX(c, nb, s, j)] : 0.0;
                    }
                }
                dgo_numerical_flux(c, c->P.conv_flux, ui, ue, nrm, fl);
                dgo_convective_flux(c, ui, F);
                for (int s = 0; s < ns; ++s) {
                    double fn = F[s * d + k];
                    if (c->curv) {
                        fn = 0.0;
                        for (int m = 0; m < d; ++m) fn += F[s * d + m] * nrm[m];
                    }
                    acc[s] += area * (fabs(fl[s]) + fabs(fn)) / c->wq[0];
                }
                const int cf = c->P.conv_flux;
                if (cf == DGO_LXF || cf == DGO_ROE || cf == DGO_TWO_POINT_LXF || cf == DGO_TWO_POINT_ROE) {
                    const double li = max_normal_eig(c, ui, nrm), le = max_normal_eig(c, ue, nrm);
                    const double lam = li > le ? li : le;
                    for (int s = 0; s < ns; ++s) acc[s] += area * 0.5 * lam * (fabs(ui[s]) + fabs(ue[s])) / c->wq[0];
                }
                if (visc) {
                    if (nb < 0) visc_boundary_state(c, -nb, ui, gi, ue, ge);
                    viscous_num_flux(c, ui, ue, gi, ge, nrm, penalty, fl, nb < 0 && wall_adiabatic(c, -nb));
                    dgo_dissipative_flux(c, ui, gi, F);
                    for (int s = 0; s < ns; ++s) {
                        double fn = F[s * d + k];
                        if (c->curv) {
                            fn = 0.0;
                            for (int m = 0; m < d; ++m) fn += F[s * d + m] * nrm[m];
                        }
                        acc[s] += area * (fabs(fl[s]) + fabs(fn)) / c->wq[0];
                    }
                    if (c->P.diss_flux == DGO_SIPG) {
                        /* penalty {{F_v(u, (u_i - u_e) n)}}: the one-sided terms in u_i and u_e */
                        const double *side_u[2] = {ui, ue};
                        for (int a = 0; a < 2; ++a)
                            for (int b = 0; b < 2; ++b) {
                                double g[NS_MAX * 3], Fv[NS_MAX * 3];
                                for (int s = 0; s < ns; ++s)
                                    for (int kk = 0; kk < d; ++kk) g[s * d + kk] = side_u[b][s] * nrm[kk];
                                dgo_dissipative_flux(c, side_u[a], g, Fv);
                                for (int s = 0; s < ns; ++s) {
                                    double fn = Fv[s * d + k];
                                    if (c->curv) {
                                        fn = 0.0;
                                        for (int m = 0; m < d; ++m) fn += Fv[s * d + m] * nrm[m];
                                    }
                                    acc[s] += area * 0.5 * penalty * fabs(fn) / c->wq[0];
                                }
                            }
                    }
                }
            }
            for (int s = 0; s < ns; ++s) mag[IDX(c, e, s, i)] = c->wvol[i] * (c->curv ? 1.0 : hd1) * acc[s];
        }
    }
}

void dgo_apply_inverse_mass(const dgo_ctx *c, const double *in, double *out, int aux)
{
    /* apply_inverse_global_mass_matrix, Cartesian branch (dg_base.cpp:4208-4219); cDG/kDG => M^-1 */
    (void)aux;
    const int ns = c->ns, n = c->n;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(g_threads)
#endif
    for (int e = 0; e < c->ncells; ++e)
        for (int s = 0; s < ns; ++s) {
            if (c->curv) {
                /* curvilinear, collocated GLL: M = W det J (diagonal; the weight-adjusted inverse,
                   dg_base.cpp:4220-4247, is the same operator here) */
                for (int i = 0; i < c->Np; ++i)
                    out[IDX(c, e, s, i)] = in[IDX(c, e, s, i)] / (c->wvol[i] * c->metJ[(long)e * c->Np + i]);
            } else {
                vol_apply(c, c->minv, n, n, in + IDX(c, e, s, 0), out + IDX(c, e, s, 0), 0, 1.0 / c->detJ);
            }
        }
}

void dgo_get_auxiliary(const dgo_ctx *c, int k, double *out)
{
    if (c->aux[k]) memcpy(out, c->aux[k], sizeof(double) * dgo_n_dofs(c));
}

/* ------------------------------------------------------------------------ */
/* initial conditions                                                       */
/* ------------------------------------------------------------------------ */
static void prim_to_cons(const dgo_ctx *c, const double *prim, double *u)
{
    /* convert_primitive_to_conservative (euler.cpp:203-218), total energy (:308-330) */
    const int d = c->dim, ns = c->ns;
    u[0] = prim[0];
    double vel2 = 0.0;
    for (int k = 0; k < d; ++k) { u[1 + k] = prim[0] * prim[1 + k]; vel2 = vel2 + prim[1 + k] * prim[1 + k]; }
    u[ns - 1] = prim[ns - 1] / c->gamm1 + 0.5 * prim[0] * vel2;
}

void dgo_ic_tgv(const dgo_ctx *c, int isothermal, double *u)
{
    /* InitialConditionFunction_TaylorGreenVortex (initial_condition_function.cpp:261-336), interpolated
       at the FE support points (set_initial_condition.cpp:45-51) */
    const int Np = c->Np, ns = c->ns;
    double *xyz = (double *)malloc(sizeof(double) * c->ncells * Np * 3);
    dgo_node_coords(c, xyz);
    for (int e = 0; e < c->ncells; ++e)
        for (int i = 0; i < Np; ++i) {
            const double x = xyz[((long)e * Np + i) * 3], y = xyz[((long)e * Np + i) * 3 + 1], z = xyz[((long)e * Np + i) * 3 + 2];
            double prim[NS_MAX];
            const double p = 1.0 / (c->gam * c->mach2) + (1.0 / 16.0) * (cos(2.0 * x) + cos(2.0 * y)) * (cos(2.0 * z) + 2.0);
            prim[0] = isothermal ? p * c->gam * c->mach2 : 1.0;
            prim[1] = sin(x) * cos(y) * cos(z);
            prim[2] = -cos(x) * sin(y) * cos(z);
            if (c->dim == 3) prim[3] = 0.0;
            prim[ns - 1] = p;
            double uu[NS_MAX];
            prim_to_cons(c, prim, uu);
            for (int s = 0; s < ns; ++s) u[IDX(c, e, s, i)] = uu[s];
        }
    free(xyz);
}

void dgo_ic_vortex(const dgo_ctx *c, double strength, double u_inf, double v_inf, double *u)
{
    /* InitialConditionFunction_IsentropicVortex (initial_condition_function.cpp:643-690) with the
       strength and advection velocity as parameters: beta = M sqrt(2) strength/(4 pi) e^(1/2);
       strength = 5 and (u,v) = M (cos pi/4, sin pi/4) is the reference's hard-coded case. */
    const int Np = c->Np, ns = c->ns;
    const double gam = 1.4;
    const double M = sqrt(2 / gam);
    const double beta = M * strength * sqrt(2.0) / 4.0 / M_PI * exp(1.0 / 2.0);
    double *xyz = (double *)malloc(sizeof(double) * c->ncells * Np * 3);
    dgo_node_coords(c, xyz);
    for (int e = 0; e < c->ncells; ++e)
        for (int i = 0; i < Np; ++i) {
            const double x = xyz[((long)e * Np + i) * 3], y = xyz[((long)e * Np + i) * 3 + 1];
            const double Omega = beta * exp(-0.5 / 1.0 / 1.0 * (x / 1.0 * x / 1.0 + y / 1.0 * y / 1.0));
            const double dux = -y / 1.0 * Omega, duy = x / 1.0 * Omega;
            const double dT = -(gam - 1.0) / 2.0 * Omega * Omega;
            double prim[NS_MAX];
            prim[0] = pow((1 + dT), 1.0 / (gam - 1.0));
            prim[1] = u_inf + dux;
            prim[2] = v_inf + duy;
            if (c->dim == 3) prim[3] = 0;
            prim[ns - 1] = 1.0 / gam * pow(1 + dT, gam / (gam - 1.0));
            double uu[NS_MAX];
            prim_to_cons(c, prim, uu);
            for (int s = 0; s < ns; ++s) u[IDX(c, e, s, i)] = uu[s];
        }
    free(xyz);
}

/* ------------------------------------------------------------------------ */
/* time integration                                                          */
/* ------------------------------------------------------------------------ */
typedef struct { int s; double a[4][4], b[4], c[4]; } butcher;

static int get_butcher(int m, butcher *B)
{
    /* runge_kutta_methods.cpp:8-110 */
    memset(B, 0, sizeof(*B));
    if (m == 0) { B->s = 1; B->b[0] = 1.0; return 1; }
    if (m == 1) { B->s = 2; B->a[1][0] = 1.0; B->b[0] = 0.5; B->b[1] = 0.5; B->c[1] = 1.0; return 1; }
    if (m == 2) {
        B->s = 3; B->a[1][0] = 1.0; B->a[2][0] = 0.25; B->a[2][1] = 0.25;
        B->b[0] = 1.0 / 6.0; B->b[1] = 1.0 / 6.0; B->b[2] = 2.0 / 3.0;
        B->c[1] = 1.0; B->c[2] = 0.5; return 1;
    }
    if (m == 3) {
        B->s = 4; B->a[1][0] = 0.5; B->a[2][1] = 0.5; B->a[3][2] = 1.0;
        B->b[0] = 1.0 / 6.0; B->b[1] = 1.0 / 3.0; B->b[2] = 1.0 / 3.0; B->b[3] = 1.0 / 6.0;
        B->c[1] = 0.5; B->c[2] = 0.5; B->c[3] = 1.0; return 1;
    }
    return 0;
}

typedef struct { int s; int plus; double gamma[11][3]; double beta[11]; double delta[10]; double bhat[11]; } lsrk;

static int get_lsrk(int m, lsrk *L)
{
    /* low_storage_runge_kutta_methods.cpp:11-220 (Ketcheson 2010; Ranocha et al. 2021) */
    memset(L, 0, sizeof(*L));
    if (m == 10) {
        static const double g[6][3] = {{0.0, 0.0, 0.0}, {0.0, 1.0, 0.0},
            {-0.497531095840104, 1.384996869124138, 0.0}, {1.010070514199942, 3.878155713328178, 0.0},
            {-3.196559004608766, -2.324512951813145, 1.642598936063715},
            {1.717835630267259, -0.514633322274467, 0.188295940828347}};
        static const double be[6] = {0.0, 0.075152045700771, 0.211361016946069, 1.100713347634329,
                                     0.728537814675568, 0.393172889823198};
        static const double de[7] = {1.0, 0.081252332929194, -1.083849060586449, -1.096110881845602,
                                     2.859440022030827, -0.655568367959557, -0.194421504490852};
        L->s = 5; L->plus = 0;
        memcpy(L->gamma, g, sizeof(g)); memcpy(L->beta, be, sizeof(be)); memcpy(L->delta, de, sizeof(de));
        return 1;
    }
    if (m == 11) {
        static const double g[6][3] = {{0.0, 0.0, 0.0}, {0.0, 1.0, 0.0},
            {0.2587771979725733308135192812685323706, 0.5528354909301389892439698870483746541, 0.0},
            {-0.1324380360140723382965420909764953437, 0.6731871608203061824849561782794643600, 0.0},
            {0.05056033948190826045833606441415585735, 0.2803103963297672407841316576323901761, 0.2752563273304676380891217287572780582},
            {0.5670532000739313812633197158607642990, 0.5521525447020610386070346724931300367, -0.8950526174674033822276061734289327568}};
        static const double be[6] = {0.0, 0.2300298624518076223899418286314123354, 0.3021434166948288809034402119555380003,
                                     0.8025606185416310937583009085873554681, 0.4362158943603440930655148245148766471, 0.11292725304550591};
        static const double de[5] = {1.0, 0.34076558793345252, 0.34143826550033862, 0.72292753667879872, 0.0};
        static const double bh[6] = {0.094841667050357029, 0.17263713394303537, 0.39982431890843712, 0.17180168075801786,
                                     0.058819144221557401, 0.1020760551185952388626787099944507877};
        L->s = 5; L->plus = 1;
        memcpy(L->gamma, g, sizeof(g)); memcpy(L->beta, be, sizeof(be)); memcpy(L->delta, de, sizeof(de)); memcpy(L->bhat, bh, sizeof(bh));
        return 1;
    }
    if (m == 12) {
        static const double g[10][3] = {{0.0, 0.0, 0.0}, {0.0, 1.0, 0.0},
            {-4.655641447335068552684422206224169103, 2.499262792574495009336242992898153462, 0.0},
            {-0.7720265099645871829248487209517314217, 0.5866820377718875577451517985847920081, 0.0},
            {-4.024436690519806086742256154738379161, 1.205146086523094569925592464380295241, 0.7621006678721315291614677352949377871},
            {-0.02129676284018530966221583708648634733, 0.3474793722186732780030762737753849272, -0.1981182504339400567765766904309673119},
            {-2.435022509790109546199372365866450709, 1.321346060965113109321230804210670518, -0.6228959218699007450469629366684127462},
            {0.01985627297131987000579523283542615256, 0.3119636464694193615946633676950358444, -0.3752248380775956442989480369774937099},
            {-0.2810791146791038566946663374735713961, 0.4351419539684379261368971206040518552, -0.3355438309135169811915662336248989661},
            {0.1689434168754859644351230590422137972, 0.2359698130028753572503744518147537768, -0.04560955005031121479972862973705108039}};
        static const double be[10] = {0.0, 0.2836343005184365275160654678626695428, 0.9736500104654741223716056170419660217, 0.3382359225242515288768487569778320563,
                                      -0.3584943611106183357043212309791897386, -0.004113944068471528211627210454497620358, 1.427968894048586363415504654313371031,
                                      0.01808470948394314017665968411915568633, 0.1605770645946802213926893453819236685, 0.2952227015964591648775833803635147962};
        static const double de[9] = {1.0, 1.262923876648114432874834923838556100, 0.7574967189685911558308119415539596711, 0.5163589453140728104667573195005629833,
                                     -0.02746327421802609557034437892013640319, -0.4382673178127944142238606608356542890, 1.273587294602656522645691372699677063,
                                     -0.6294740283927400326554066998751383342, 0.0};
        static const double bh[10] = {0.02483675912451591196775756814283216443, 0.1866327774562103796990092260942180726, 0.05671080795936984495604436622517631183,
                                      -0.003447695439149287702616943808570747099, 0.003602245056516636472203469198006404016, 0.4545570622145088936800484247980581766,
                                      -0.0002434665289427612407531544765622888855, 0.06642755361103549971517945063138312147, 0.1613697079523505006226025497715177578,
                                      0.04955424859358438183052504342394102722};
        L->s = 9; L->plus = 1;
        memcpy(L->gamma, g, sizeof(g)); memcpy(L->beta, be, sizeof(be)); memcpy(L->delta, de, sizeof(de)); memcpy(L->bhat, bh, sizeof(bh));
        return 1;
    }
    if (m == 13) {
        static const double g[11][3] = {{0.0, 0.0, 0.0}, {0.0, 1.0, 0.0},
            {0.4043660121685749695640462197806189975, 0.6871467028161416909922221357014564412, 0.0},
            {-0.8503427289575839690883191973980814832, 1.093024748914750833700799552463885117, 0.0},
            {-6.950894175262117526410215315179482885, 3.225975379607193001678365742708874597, -2.393405133244194727221124311276648940},
            {0.9238765192731084931855438934978371889, 1.041153702510101386914019859778740444, -1.902854422421760920850597670305403139},
            {-2.563178056509891340215942413817786020, 1.292821487912164945157744726076279306, -2.820042207399977261483046412236557428},
            {0.2545744879365226143946122067064118430, 0.7391462755788122847651304143259254381, -1.832698465277380999601896111079977378},
            {0.3125831707411998258746812355492206137, 0.1239129251371800313941948224441873274, -0.2199094483084671192328083958346519535},
            {-0.7007114414440507927791249989236719346, 0.1842753472370123193132193302369345580, -0.4082430635847870963724591602173546218},
            {0.4839621016023833375810172323297465039, 0.05712788998796583446479387686662738843, -0.1377669797880289713535665985132703979}};
        static const double be[11] = {0.0, 0.2597883554788674084039539165398464630, 0.01777008889438867858759149597539211023, 0.2481636629715501931294746189266601496,
                                      0.7941736871152005775821844297293296135, 0.3885391285642019129575902994397298066, 0.1455051657916305055730603387469193768,
                                      0.1587517385964749337690916959584348979, 0.1650605617880053419242434594242509601, 0.2118093284937153836908655490906875007,
                                      0.1559392342362059886106995325687547506};
        static const double de[10] = {1.0, -0.1331778419508803397033287009506932673, 0.8260422814750207498262063505871077303,
                                      1.513700425755728332485300719652378197, -1.305810059935023735972298885749903694, 3.036678802924163246003321318996156380,
                                      -1.449458274398895177922690618003584514, 3.834313899176362315089976408899373409, 4.122293760012985409330881631526514714,
                                      0.0};
        static const double bh[11] = {-0.02019255440012066080909442770590267512, 0.02737903480959184339932730854141598275, 0.3028818636145965534365173822296811090,
                                      -0.03656843880622222190071445247906780540, 0.3982664774676767729863101188528827405, -0.05715959421140685436681459970502471634,
                                      0.09849855103848558320961101178888983150, 0.06654601552456084978615342374581437947, 0.09073479542748112726465375642050504556,
                                      0.08432289325330803924891866923939606351, 0.04529095628204896774513180907141004447};
        L->s = 10; L->plus = 1;
        memcpy(L->gamma, g, sizeof(g)); memcpy(L->beta, be, sizeof(be)); memcpy(L->delta, de, sizeof(de)); memcpy(L->bhat, bh, sizeof(bh));
        return 1;
    }
    return 0;
}

int dgo_n_stages(int m)
{
    butcher B;
    lsrk L;
    if (get_butcher(m, &B)) return B.s;
    if (get_lsrk(m, &L)) return L.s;
    return 0;
}

static void rhs_lifted(dgo_ctx *c, const double *u, double *k, double *scratch)
{
    dgo_assemble_residual(c, u, scratch);
    dgo_apply_inverse_mass(c, scratch, k, 0);
}

/* LowStorageRungeKuttaODESolver (low_storage_runge_kutta_ode_solver.cpp:24-79, :207-222) one
   step; with w != NULL also the error norm of the embedded pair that
   get_automatic_error_adaptive_step_size (:89-118) computes after the step: S1 against S4
   (3S*+, FSAL residual evaluated) or against the embedded solution S2 of sum_stages (:55-62). */
static void lsrk_step(dgo_ctx *c, int m, const lsrk *Lp, double dt, double *u, double *scratch, double atol,
                      double rtol, double *w)
{
    const lsrk L = *Lp;
    const long nd = dgo_n_dofs(c);
    double *S1 = (double *)malloc(sizeof(double) * nd);
    double *S2 = (double *)calloc((size_t)nd, sizeof(double));
    double *S3 = (double *)malloc(sizeof(double) * nd);
    double *S4 = (double *)malloc(sizeof(double) * nd);
    double *r = (double *)malloc(sizeof(double) * nd);
    memcpy(S1, u, sizeof(double) * nd);
    memcpy(S3, S1, sizeof(double) * nd);
    memcpy(r, S1, sizeof(double) * nd);
    memcpy(S4, S1, sizeof(double) * nd);
    for (int i = 0; i < L.s; ++i) {
        for (long q = 0; q < nd; ++q) S2[q] += L.delta[i] * S1[q];
        memcpy(u, r, sizeof(double) * nd);
        rhs_lifted(c, u, r, scratch);
        for (long q = 0; q < nd; ++q) {
            S1[q] *= L.gamma[i + 1][0];
            S1[q] += L.gamma[i + 1][1] * S2[q];
            S1[q] += L.gamma[i + 1][2] * S3[q];
            r[q] *= dt;
            S1[q] += L.beta[i + 1] * r[q];
            if (L.plus) S4[q] += L.bhat[i] * r[q];
        }
        memcpy(r, S1, sizeof(double) * nd);
    }
    if (L.plus && (c->fsal || w)) {
        /* FSAL residual for the embedded estimate (:63-69); it only updates S4 */
        memcpy(u, r, sizeof(double) * nd);
        rhs_lifted(c, u, r, scratch);
        for (long q = 0; q < nd; ++q) { r[q] *= dt; S4[q] += L.bhat[L.s] * r[q]; }
    }
    if (!L.plus) {
        /* embedded solution (:55-62): S2 = (S2 + delta_s S1 + delta_{s+1} S3) / sum of the num_delta deltas */
        const int num_delta = (m == 10) ? 7 : L.s;
        double sum_delta = 0.0;
        for (int i = 0; i < num_delta; ++i) sum_delta += L.delta[i];
        for (long q = 0; q < nd; ++q) {
            S2[q] += L.delta[L.s] * S1[q];
            S2[q] += L.delta[L.s + 1] * S3[q];
            S2[q] /= sum_delta;
        }
    }
    if (w) {
        const double *E = L.plus ? S4 : S2;
        double acc = 0.0;
        for (long q = 0; q < nd; ++q) {
            const double err = S1[q] - E[q];
            const double a = fabs(S1[q]), b = fabs(E[q]);
            acc = acc + pow(err / (atol + rtol * (a > b ? a : b)), 2);
        }
        *w = pow(acc / (double)nd, 0.5);
    }
    memcpy(u, S1, sizeof(double) * nd);
    free(S1); free(S2); free(S3); free(S4); free(r);
}

double dgo_step_error_norm(dgo_ctx *c, int m, double dt, double *u, double atol, double rtol)
{
    lsrk L;
    if (!get_lsrk(m, &L)) return -1.0;
    double *scratch = (double *)malloc(sizeof(double) * dgo_n_dofs(c));
    double w = 0.0;
    lsrk_step(c, m, &L, dt, u, scratch, atol, rtol, &w);
    free(scratch);
    return w;
}

/* LowStorageRungeKuttaODESolver::get_automatic_initial_step_size (low_storage_runge_kutta_ode_solver.cpp:119-178) */
double dgo_initial_step_size(dgo_ctx *c, const double *u)
{
    const long nd = dgo_n_dofs(c);
    double *un = (double *)malloc(sizeof(double) * nd);
    double *f0 = (double *)malloc(sizeof(double) * nd);
    double *f1 = (double *)malloc(sizeof(double) * nd);
    double *scratch = (double *)malloc(sizeof(double) * nd);
    double d0 = 0.0, d1 = 0.0, d2 = 0.0, h0, h1;
    for (long q = 0; q < nd; ++q) d0 = fabs(u[q]) > d0 ? fabs(u[q]) : d0;
    rhs_lifted(c, u, f0, scratch);
    for (long q = 0; q < nd; ++q) d1 = fabs(f0[q]) > d1 ? fabs(f0[q]) : d1;
    h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    for (long q = 0; q < nd; ++q) un[q] = u[q] + h0 * f0[q];
    rhs_lifted(c, un, f1, scratch);
    for (long q = 0; q < nd; ++q) {
        const double x = fabs(f1[q] - f0[q]);
        d2 = x > d2 ? x : d2;
    }
    d2 /= h0;
    const double dm = d1 > d2 ? d1 : d2;
    if (dm <= 1e-15) h1 = (1e-6 > h0 * 1e-3) ? 1e-6 : h0 * 1e-3;
    else h1 = 0.01 / dm;
    free(un); free(f0); free(f1); free(scratch);
    return (100 * h0 < h1) ? 100 * h0 : h1;
}

void dgo_step(dgo_ctx *c, int m, double dt, double *u)
{
    const long nd = dgo_n_dofs(c);
    butcher B;
    lsrk L;
    double *scratch = (double *)malloc(sizeof(double) * nd);
    if (get_butcher(m, &B)) {
        /* RungeKuttaBase::step_in_time + RungeKuttaODESolver (runge_kutta_base.cpp:21-38,
           runge_kutta_ode_solver.cpp:15-99) */
        double *upd = (double *)malloc(sizeof(double) * nd);
        double *stage[4];
        memcpy(upd, u, sizeof(double) * nd);
        for (int i = 0; i < B.s; ++i) {
            stage[i] = (double *)calloc((size_t)nd, sizeof(double));
            for (int j = 0; j < i; ++j)
                if (B.a[i][j] != 0)
                    for (long q = 0; q < nd; ++q) stage[i][q] += B.a[i][j] * stage[j][q];
            for (long q = 0; q < nd; ++q) stage[i][q] *= dt;
            for (long q = 0; q < nd; ++q) stage[i][q] += 1.0 * upd[q];
            memcpy(u, stage[i], sizeof(double) * nd); /* dg->solution = rk_stage[i] */
            rhs_lifted(c, u, stage[i], scratch);
        }
        for (int i = 0; i < B.s; ++i)
            for (long q = 0; q < nd; ++q) upd[q] += dt * B.b[i] * stage[i][q];
        memcpy(u, upd, sizeof(double) * nd);
        for (int i = 0; i < B.s; ++i) free(stage[i]);
        free(upd);
    } else if (get_lsrk(m, &L)) {
        lsrk_step(c, m, &L, dt, u, scratch, 0.0, 0.0, NULL);
    }
    free(scratch);
}

double dgo_constant_time_step(const dgo_ctx *c, double cfl)
{
    /* PeriodicTurbulence::get_constant_time_step (periodic_turbulence.cpp:115-125) */
    const unsigned ndofs_per_state = (unsigned)(dgo_n_dofs(c) / c->ns);
    const double dx = (c->P.domain_right - c->P.domain_left) / pow(ndofs_per_state, (1.0 / c->dim));
    return cfl * dx;
}

double dgo_vortex_time_step(const dgo_ctx *c, double cfl, double speed)
{
    /* PeriodicEntropyTests::get_constant_time_step (periodic_entropy_tests.cpp:22-36) */
    const unsigned ndofs_per_state = (unsigned)(dgo_n_dofs(c) / c->ns);
    const double dx = (c->P.domain_right - c->P.domain_left) / pow(ndofs_per_state, (1.0 / c->dim));
    return cfl * dx / speed;
}

/* ------------------------------------------------------------------------ */
/* integrated quantities (periodic_turbulence.cpp:439-671)                   */
/* ------------------------------------------------------------------------ */
/* The metric terms of curvilinear cell e at the QGauss points (metric_operators::
   build_volume_metric_operators with the points as flux nodes, operators.cpp:2520-2884, as
   periodic_turbulence.cpp:487-492 builds them): the Jacobian of the interpolated mapping at the points
   (B, G: 1D values and derivatives of the grid-node Lagrange basis there), det J, and the cofactor --
   in 2D the adjugate of that Jacobian, in 3D the conservative curl form: the grid-node products
   X_l dX_m/dxi_j (their derivatives taken with the nodal derivative matrix) differentiated through
   their interpolant at the points.  Cq[(a*3+b)*nqv + q], Jq[q]. */
static void metrics_at_points(const dgo_ctx *c, int e, const double *B, const double *G, int nx, double *Cq,
                              double *Jq)
{
    const int d = c->dim, n = c->n, Np = c->Np;
    const int nqv = (d == 3) ? nx * nx * nx : nx * nx;
    double *X = (double *)malloc(sizeof(double) * 3 * Np);
    double *Jm = (double *)malloc(sizeof(double) * 9 * Np);
    double *P = (double *)malloc(sizeof(double) * 9 * Np);
    double *Jp = (double *)malloc(sizeof(double) * 9 * nqv);
    for (int i = 0; i < d; ++i)
        for (int q = 0; q < Np; ++q) X[i * Np + q] = c->xn[((long)e * Np + q) * 3 + i];
    /* the Jacobian at the points: dX_i/dxi_j of the interpolant */
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j)
            dgo_tensor_apply(d, j == 0 ? G : B, nx, n, j == 1 ? G : B, nx, n, j == 2 ? G : B, nx, n, X + i * Np,
                             Jp + (i * 3 + j) * nqv, 0, 1.0);
    memset(Cq, 0, sizeof(double) * 9 * nqv);
    if (d == 2) {
        for (int q = 0; q < nqv; ++q) {
            Cq[(0 * 3 + 0) * nqv + q] = Jp[(1 * 3 + 1) * nqv + q];
            Cq[(1 * 3 + 0) * nqv + q] = -Jp[(0 * 3 + 1) * nqv + q];
            Cq[(0 * 3 + 1) * nqv + q] = -Jp[(1 * 3 + 0) * nqv + q];
            Cq[(1 * 3 + 1) * nqv + q] = Jp[(0 * 3 + 0) * nqv + q];
        }
    } else {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) dir_apply(c, j, c->chi_grad, c->chi, n, n, X + i * Np, Jm + (i * 3 + j) * Np, 0, 1.0);
        const int lm[3][2] = {{2, 1}, {0, 2}, {1, 0}};
        for (int r = 0; r < 3; ++r)
            for (int j = 0; j < 3; ++j)
                for (int q = 0; q < Np; ++q)
                    P[(r * 3 + j) * Np + q] = X[lm[r][0] * Np + q] * Jm[(lm[r][1] * 3 + j) * Np + q];
        /* directional derivative of a grid-node field at the points */
        #define DPT(dir, in, out, add, f) \
            dgo_tensor_apply(3, (dir) == 0 ? G : B, nx, n, (dir) == 1 ? G : B, nx, n, (dir) == 2 ? G : B, nx, n, \
                             (in), (out), (add), (f))
        for (int r = 0; r < 3; ++r) {
            const double *Pxi = P + (r * 3 + 0) * Np, *Peta = P + (r * 3 + 1) * Np, *Pzeta = P + (r * 3 + 2) * Np;
            double *C0 = Cq + (r * 3 + 0) * nqv, *C1 = Cq + (r * 3 + 1) * nqv, *C2 = Cq + (r * 3 + 2) * nqv;
            DPT(1, Pzeta, C0, 0, -1.0);
            DPT(2, Peta, C0, 1, 1.0);
            DPT(0, Pzeta, C1, 0, 1.0);
            DPT(2, Pxi, C1, 1, -1.0);
            DPT(0, Peta, C2, 0, -1.0);
            DPT(1, Pxi, C2, 1, 1.0);
        }
        #undef DPT
    }
    for (int q = 0; q < nqv; ++q) {
        double J[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J[i][j] = (i < d && j < d) ? Jp[(i * 3 + j) * nqv + q] : 0.0;
        if (d == 2) Jq[q] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        else Jq[q] = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                   + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        if (g_mutation == DGO_MUT_TRANSPOSED_METRIC)
            /* planted: the cofactor transposed in the gradient transform */
            for (int a = 0; a < 3; ++a)
                for (int b = a + 1; b < 3; ++b) {
                    const double t = Cq[(a * 3 + b) * nqv + q];
                    Cq[(a * 3 + b) * nqv + q] = Cq[(b * 3 + a) * nqv + q];
                    Cq[(b * 3 + a) * nqv + q] = t;
                }
        if (g_mutation == DGO_MUT_CELL_CONSTANT_DETJ) Jq[q] = c->detJ;  /* planted: the box's det J */
    }
    free(X); free(Jm); free(P); free(Jp);
}

void dgo_integrated_quantities(const dgo_ctx *c, const double *u, int overint, double *out)
{
    const int d = c->dim, ns = c->ns, n = c->n, Np = c->Np;
    const int nx = c->p + 1 + overint;
    double xq[64], wq[64];
    dgo_gauss_legendre(nx, xq, wq);
    double *B = (double *)malloc(sizeof(double) * nx * n);
    double *G = (double *)malloc(sizeof(double) * nx * n);
    double val[MAXN], der[MAXN];
    for (int q = 0; q < nx; ++q) {
        dgo_lagrange_values(n, c->xs, xq[q], val, der);
        for (int j = 0; j < n; ++j) { B[q * n + j] = val[j]; G[q * n + j] = der[j]; }
    }
    const int nqv = (d == 3) ? nx * nx * nx : nx * nx;
    double *W = (double *)malloc(sizeof(double) * nqv);
    for (int q = 0; q < nqv; ++q) {
        const int a = q % nx, b = (q / nx) % nx, cc = (d == 3) ? q / (nx * nx) : 0;
        W[q] = (d == 3) ? wq[a] * wq[b] * wq[cc] : wq[a] * wq[b];
    }
    double ke = 0.0, ens = 0.0, ike = 0.0;
    double *sq = (double *)malloc(sizeof(double) * ns * nqv);
    double *gq = (double *)malloc(sizeof(double) * ns * d * nqv);
    double *coef = (double *)malloc(sizeof(double) * Np);
    double C[3][3];
    metric_cofactor(c, C);
    double *Cq = c->curv ? (double *)malloc(sizeof(double) * 9 * nqv) : NULL;
    double *Jq = c->curv ? (double *)malloc(sizeof(double) * nqv) : NULL;
    for (int e = 0; e < c->ncells; ++e) {
        if (c->curv) metrics_at_points(c, e, B, G, nx, Cq, Jq);
        for (int s = 0; s < ns; ++s) {
            for (int i = 0; i < Np; ++i) coef[i] = u[IDX(c, e, s, i)];
            dgo_tensor_apply(d, B, nx, n, B, nx, n, B, nx, n, coef, sq + s * nqv, 0, 1.0);
            for (int k = 0; k < d; ++k)
                dgo_tensor_apply(d, k == 0 ? G : B, nx, n, k == 1 ? G : B, nx, n, k == 2 ? G : B, nx, n,
                                 coef, gq + (s * d + k) * nqv, 0, 1.0);
        }
        for (int q = 0; q < nqv; ++q) {
            double st[NS_MAX], g[NS_MAX * 3];
            const double jq = c->curv ? Jq[q] : c->detJ;
            if (c->curv)
                for (int a = 0; a < d; ++a)
                    for (int b = 0; b < d; ++b) C[a][b] = Cq[(a * 3 + b) * nqv + q];
            for (int s = 0; s < ns; ++s) {
                st[s] = sq[s * nqv + q];
                for (int k = 0; k < d; ++k) {
                    double acc = 0.0;
                    for (int j = 0; j < d; ++j) acc += C[k][j] * gq[(s * d + j) * nqv + q] / jq;
                    g[s * d + k] = acc;
                }
            }
            double vel[3], vel2 = 0.0;
            velocities(c, st, vel);
            for (int k = 0; k < d; ++k) vel2 = vel2 + vel[k] * vel[k];
            const double KE = 0.5 * st[0] * vel2;
            /* vorticity (navier_stokes.cpp:396-423) from primitive velocity gradient */
            double vg[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            for (int a = 0; a < d; ++a)
                for (int b = 0; b < d; ++b) vg[a][b] = (g[(1 + a) * d + b] - vel[a] * g[0 * d + b]) / st[0];
            double om[3] = {0, 0, 0};
            if (d == 2) om[2] = vg[1][0] - vg[0][1];
            else { om[0] = vg[2][1] - vg[1][2]; om[1] = vg[0][2] - vg[2][0]; om[2] = vg[1][0] - vg[0][1]; }
            double om2 = 0.0;
            for (int k = 0; k < 3; ++k) om2 += om[k] * om[k];
            const double ENS = 0.5 * st[0] * om2;
            const double w = W[q] * jq;
            ke += KE * w;
            ens += ENS * w;
            ike += 0.5 * vel2 * w;
        }
    }
    double vol = 1.0;
    for (int k = 0; k < d; ++k) vol *= (c->P.domain_right - c->P.domain_left);
    out[0] = ke / vol;
    out[1] = ens / vol;
    out[2] = ike / vol;
    out[3] = 2.0 * out[1] / c->P.reynolds_inf;
    free(B); free(G); free(W); free(sq); free(gq); free(coef); free(Cq); free(Jq);
}

double dgo_entropy_rate(dgo_ctx *c, const double *u)
{
    /* semi-discrete entropy production  sum_e v^T M (M^-1 R) = sum_e v^T R(u), with v the entropy
       variables at the solution nodes (collocated GLL: projection is the identity) */
    const long nd = dgo_n_dofs(c);
    double *r = (double *)malloc(sizeof(double) * nd);
    dgo_assemble_residual(c, u, r);
    double acc = 0.0;
    for (int e = 0; e < c->ncells; ++e)
        for (int i = 0; i < c->Np; ++i) {
            double us[NS_MAX], vs[NS_MAX];
            for (int s = 0; s < c->ns; ++s) us[s] = u[IDX(c, e, s, i)];
            dgo_entropy_variables(c, us, vs);
            for (int s = 0; s < c->ns; ++s) acc += vs[s] * r[IDX(c, e, s, i)];
        }
    free(r);
    return acc;
}

/* PeriodicTurbulence::update_maximum_local_wave_speed (periodic_turbulence.cpp:371-422):
   max over cells and QGauss(p+1+overint) points of |u| + c (Euler::max_convective_eigenvalue,
   euler.cpp:1033-1046), the solution interpolated from the GLL support points. */
double dgo_max_wave_speed(const dgo_ctx *c, const double *u, int overint)
{
    const int d = c->dim, ns = c->ns, n = c->n, Np = c->Np;
    const int nx = c->p + 1 + overint;
    double xq[64], wq[64];
    dgo_gauss_legendre(nx, xq, wq);
    double *B = (double *)malloc(sizeof(double) * nx * n);
    double val[MAXN], der[MAXN];
    for (int q = 0; q < nx; ++q) {
        dgo_lagrange_values(n, c->xs, xq[q], val, der);
        for (int j = 0; j < n; ++j) B[q * n + j] = val[j];
    }
    const int nqv = (d == 3) ? nx * nx * nx : nx * nx;
    double *sq = (double *)malloc(sizeof(double) * ns * nqv);
    double *coef = (double *)malloc(sizeof(double) * Np);
    double vmax = 0.0;
    for (int e = 0; e < c->ncells; ++e) {
        for (int st = 0; st < ns; ++st) {
            for (int i = 0; i < Np; ++i) coef[i] = u[IDX(c, e, st, i)];
            dgo_tensor_apply(d, B, nx, n, B, nx, n, B, nx, n, coef, sq + st * nqv, 0, 1.0);
        }
        for (int q = 0; q < nqv; ++q) {
            double us[NS_MAX];
            for (int st = 0; st < ns; ++st) us[st] = sq[st * nqv + q];
            const double ev = dgo_max_convective_eigenvalue(c, us);
            if (ev > vmax) vmax = ev;
        }
    }
    free(B); free(sq); free(coef);
    return vmax;
}

/* PeriodicTurbulence::get_adaptive_time_step (periodic_turbulence.cpp:359-367) */
double dgo_adaptive_time_step(const dgo_ctx *c, const double *u, double cfl)
{
    const unsigned ndofs_per_state = (unsigned)(dgo_n_dofs(c) / c->ns);
    const double dx = (c->P.domain_right - c->P.domain_left) / pow(ndofs_per_state, (1.0 / c->dim));
    return cfl * dx / dgo_max_wave_speed(c, u, 10);
}

/* PeriodicTurbulence::compute_current_integrated_numerical_entropy
   (periodic_turbulence.cpp:765-868): sum over cells and volume quadrature points of
   w_q det J * (-rho s), s = ln(p rho^-gamma) (euler.cpp:437-447, 511-532), the solution
   interpolated to the volume quadrature (identity for collocated GLL). */
double dgo_numerical_entropy(const dgo_ctx *c, const double *u)
{
    const int ns = c->ns, Np = c->Np, Nq = (c->dim == 3) ? c->nq * c->nq * c->nq : c->nq * c->nq;
    double *uq = (double *)malloc(sizeof(double) * ns * Nq);
    double acc = 0.0;
    for (int e = 0; e < c->ncells; ++e) {
        for (int st = 0; st < ns; ++st) vol_apply(c, c->chi, c->nq, c->n, u + IDX(c, e, st, 0), uq + st * Nq, 0, 1.0);
        for (int q = 0; q < Nq; ++q) {
            double us[NS_MAX];
            for (int st = 0; st < ns; ++st) us[st] = uq[st * Nq + q];
            const double rho = us[0];
            const double p = dgo_pressure(c, us);
            const double ent = log(p * pow(rho, -c->gam));
            acc += (-rho * ent) * c->wvol[q] * jac_at(c, e, q);
        }
    }
    free(uq);
    (void)Np;
    return acc;
}

/* Relaxation Runge-Kutta, root-finding form (RRKODESolverBase::update_relaxation_parameter,
   rrk_ode_solver_base.cpp:16-35; RootFindingRRKODESolver::compute_relaxation_parameter,
   root_finding_rrk_ode_solver.cpp:16-176 with the secant iteration and its bisection fallback;
   compute_entropy_change_estimate, :357-390; RungeKuttaBase::step_in_time, runge_kutta_base.cpp:21-38):
   one Butcher step of size dt whose update u_n + gamma dt sum b_i k_i makes the integrated
   numerical entropy change equal gamma times the estimate dt sum b_i v(u^(i)) . M k_i.
   *gamma: in, the previous step's relaxation parameter (1 initially); out, this step's.
   Returns the time advanced, gamma dt (-1 if the root finding failed). */
static double rrk_root_function(dgo_ctx *c, double gamma, const double *un, const double *dir, double s_n,
                                double est, double *tmp)
{
    const long nd = dgo_n_dofs(c);
    for (long q = 0; q < nd; ++q) tmp[q] = un[q] + gamma * dir[q];
    return dgo_numerical_entropy(c, tmp) - s_n - gamma * est;
}

double dgo_step_rrk(dgo_ctx *c, int m, double dt, double *u, double *gamma_io, double tol)
{
    butcher B;
    if (!get_butcher(m, &B)) return -1.0;
    const long nd = dgo_n_dofs(c);
    const int d = c->dim, ns = c->ns, Np = c->Np;
    double *un = (double *)malloc(sizeof(double) * nd);
    double *scratch = (double *)malloc(sizeof(double) * nd);
    double *tmp = (double *)malloc(sizeof(double) * nd);
    double *dir = (double *)calloc((size_t)nd, sizeof(double));
    double *stage[4], *stage_sol[4];
    memcpy(un, u, sizeof(double) * nd);
    double est = 0.0;
    for (int i = 0; i < B.s; ++i) {
        stage[i] = (double *)calloc((size_t)nd, sizeof(double));
        stage_sol[i] = (double *)malloc(sizeof(double) * nd);
        for (int j = 0; j < i; ++j)
            if (B.a[i][j] != 0)
                for (long q = 0; q < nd; ++q) stage[i][q] += B.a[i][j] * stage[j][q];
        for (long q = 0; q < nd; ++q) stage[i][q] *= dt;
        for (long q = 0; q < nd; ++q) stage[i][q] += 1.0 * un[q];
        memcpy(stage_sol[i], stage[i], sizeof(double) * nd);
        rhs_lifted(c, stage_sol[i], stage[i], scratch);
        /* entropy variables of the stage solution . M k_i (collocated GLL: the projection is the
           identity, M = W det J) */
        double ent = 0.0;
        for (int e = 0; e < c->ncells; ++e)
            for (int q = 0; q < Np; ++q) {
                double us[NS_MAX], vs[NS_MAX];
                for (int st = 0; st < ns; ++st) us[st] = stage_sol[i][IDX(c, e, st, q)];
                dgo_entropy_variables(c, us, vs);
                for (int st = 0; st < ns; ++st) ent += vs[st] * (c->wvol[q] * jac_at(c, e, q) * stage[i][IDX(c, e, st, q)]);
            }
        est += B.b[i] * ent;
    }
    est *= dt;
    for (int i = 0; i < B.s; ++i)
        for (long q = 0; q < nd; ++q) dir[q] += B.b[i] * stage[i][q];
    for (long q = 0; q < nd; ++q) dir[q] *= dt;
    const double s_n = dgo_numerical_entropy(c, un);
    /* secant from the previous relaxation parameter +- 1e-5, bisection if it fails */
    double gamma_kp1 = 0.0;
    int iter = 0, failed = 0;
    const int limit = 100;
    {
        double residual = 1.0;
        double gk = *gamma_io + 1e-5, gkm1 = *gamma_io - 1e-5;
        double rk = rrk_root_function(c, gk, un, dir, s_n, est, tmp);
        double rkm1 = rrk_root_function(c, gkm1, un, dir, s_n, est, tmp);
        while (residual > tol && iter < limit) {
            if (rkm1 == rk) {
                gk *= 1.001;
                rkm1 = rrk_root_function(c, gkm1, un, dir, s_n, est, tmp);
                rk = rrk_root_function(c, gk, un, dir, s_n, est, tmp);
            }
            if (gk < 0.5 || gk > 1.5) {
                gk = 1.0;
                rk = rrk_root_function(c, gk, un, dir, s_n, est, tmp);
            }
            gamma_kp1 = gk - rk * (gk - gkm1) / (rk - rkm1);
            residual = fabs(gamma_kp1 - gk);
            iter++;
            gkm1 = gk;
            gk = gamma_kp1;
            rkm1 = rk;
            rk = rrk_root_function(c, gk, un, dir, s_n, est, tmp);
            if (isnan(gk) || isnan(gkm1)) {
                gk = 1.0 - 1e-5;
                rk = rrk_root_function(c, gk, un, dir, s_n, est, tmp);
                gkm1 = 1.0 + 1e-5;
                rkm1 = rrk_root_function(c, gkm1, un, dir, s_n, est, tmp);
                residual = 1.0;
            }
        }
        if (iter == limit) failed = 1;
    }
    if (failed) {
        iter = 0;
        double lo = *gamma_io - 0.1, hi = *gamma_io + 0.1;
        double rlo = rrk_root_function(c, lo, un, dir, s_n, est, tmp);
        double rhi = rrk_root_function(c, hi, un, dir, s_n, est, tmp);
        double residual = 1.0;
        while (residual > tol && iter < limit) {
            if (rlo * rhi > 0) {
                lo -= 0.1;
                hi += 0.1;
            }
            gamma_kp1 = 0.5 * (lo + hi);
            const double r = rrk_root_function(c, gamma_kp1, un, dir, s_n, est, tmp);
            if (r < 0) { lo = gamma_kp1; rlo = r; }
            else { hi = gamma_kp1; rhi = r; }
            residual = hi - lo;
            iter++;
        }
    }
    double advanced = -1.0;
    if (iter < limit && gamma_kp1 >= 0.5) {
        /* sum_stages with the relaxed step gamma dt (runge_kutta_ode_solver.cpp:86-98) */
        const double dtr = dt * gamma_kp1;
        memcpy(u, un, sizeof(double) * nd);
        for (int i = 0; i < B.s; ++i)
            for (long q = 0; q < nd; ++q) u[q] += dtr * B.b[i] * stage[i][q];
        *gamma_io = gamma_kp1;
        advanced = dtr;
    }
    for (int i = 0; i < B.s; ++i) { free(stage[i]); free(stage_sol[i]); }
    free(un); free(scratch); free(tmp); free(dir);
    (void)d;
    return advanced;
}

/* FlowSolver::run time loop (flow_solver.cpp:554-610) with constant dt */
int dgo_run(dgo_ctx *c, int method, double dt_const, double final_time, double *u, double *ke_hist, int max_hist)
{
    double t = 0.0, next_dt = dt_const;
    int it = 0;
    while (t < final_time) {
        double dt = next_dt;
        if ((t + dt) > final_time) dt = final_time - t;
        dgo_step(c, method, dt, u);
        t += dt;
        if (ke_hist && it < max_hist) {
            double q[4];
            dgo_integrated_quantities(c, u, 10, q);
            ke_hist[it] = q[0];
        }
        ++it;
        next_dt = dt_const;
    }
    return it;
}
