/*
 * lumen_oracle_bwd.c -- FP64 forward and analytic backward of the splat path.
 *
 * TEST INFRASTRUCTURE ONLY (see lumen_oracle.h).  The reference has no
 * backward pass (its trainer injects dummy gradients, src/optim/trainer.mojo:
 * 59-65); the gradient layout it defines is GaussianGradients
 * (src/optim/gradients.mojo:10-106): per-Gaussian planes for pos xyz,
 * scale xyz, rot wxyz, opacity and sh[48], plus viewspace_grad_2d and
 * observation_count.
 *
 * orc64_forward renders the path of src/raster/pipeline.mojo:48-231 in
 * double precision (same culls, same tile lists in (tile, depth_mm, visible
 * index) order, same alpha/skip/termination rules, one normalization where the
 * FP32 path normalizes twice -- identical in exact arithmetic), and returns a
 * signature of every discrete decision so that finite differences that cross
 * a threshold can be recognised.  orc64_backward is its exact derivative.
 * Pinned by central finite differences of orc64_forward (tests/).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lumen_oracle.h"

#define MIN_ALPHA (double)0.00392156862f
#define T_EPS (double)0.0001f

static const double SHC0 = 0.28209479177387814, SHC1 = 0.4886025119029199,
                    SHC20 = 1.0925484305920792, SHC21 = 0.31539156525252005,
                    SHC22 = 0.5462742152960396, SHC30 = 0.5900435899266435,
                    SHC31 = 2.890611442640554, SHC32 = 0.4570457994644658,
                    SHC33 = 0.3731763325901154, SHC34 = 1.445305721320277;

typedef struct {
    int64_t orig;
    double mx, my, z, a, b, c, op, col[3], colraw[3];
    int64_t rect[4];
    uint64_t key;
    /* saved for the backward */
    double pc[3], t0[3], t1[3], S[6], R[9], s[3], qhat[4], qnorm, c2[3], det;
    double dir[3], dirlen, Y[16];
    int sclamp[3], inv_zero;
} vis64;

typedef struct {
    int64_t V, tiles_x, tiles_y, W, H;
    vis64 *vis;
    int64_t *order;   /* visible indices sorted by (depth key, index) */
    int64_t *tstart;  /* per tile: offset into tlist */
    int64_t *tlist;   /* concatenated per-tile visible indices in composite order */
    double W9[9];
} scene_state;

static void mat_from_quat(const double q[4], double R[9]) {
    double w = q[0], x = q[1], y = q[2], z = q[3];
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z); R[2] = 2 * (x * z + w * y);
    R[3] = 2 * (x * y + w * z); R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
    R[6] = 2 * (x * z - w * y); R[7] = 2 * (y * z + w * x); R[8] = 1 - 2 * (x * x + y * y);
}

static void sh_basis64(const double d[3], double Y[16]) {
    double x = d[0], y = d[1], z = d[2], xx = x * x, yy = y * y, zz = z * z;
    Y[0] = SHC0; Y[1] = -SHC1 * y; Y[2] = SHC1 * z; Y[3] = -SHC1 * x;
    Y[4] = SHC20 * x * y; Y[5] = -SHC20 * y * z; Y[6] = SHC21 * (2 * zz - xx - yy);
    Y[7] = -SHC20 * x * z; Y[8] = SHC22 * (xx - yy);
    Y[9] = -SHC30 * y * (3 * xx - yy); Y[10] = SHC31 * x * y * z;
    Y[11] = -SHC32 * y * (4 * zz - xx - yy); Y[12] = SHC33 * z * (2 * zz - 3 * xx - 3 * yy);
    Y[13] = -SHC32 * x * (4 * zz - xx - yy); Y[14] = SHC34 * z * (xx - yy);
    Y[15] = -SHC30 * x * (xx - 3 * yy);
}

/* dY_k/d(x, y, z) of the polynomial basis above. */
static void sh_basis_grad64(const double d[3], double G[16][3]) {
    double x = d[0], y = d[1], z = d[2], xx = x * x, yy = y * y, zz = z * z;
    memset(G, 0, sizeof(double) * 48);
    G[1][1] = -SHC1; G[2][2] = SHC1; G[3][0] = -SHC1;
    G[4][0] = SHC20 * y; G[4][1] = SHC20 * x;
    G[5][1] = -SHC20 * z; G[5][2] = -SHC20 * y;
    G[6][0] = -2 * SHC21 * x; G[6][1] = -2 * SHC21 * y; G[6][2] = 4 * SHC21 * z;
    G[7][0] = -SHC20 * z; G[7][2] = -SHC20 * x;
    G[8][0] = 2 * SHC22 * x; G[8][1] = -2 * SHC22 * y;
    G[9][0] = -6 * SHC30 * x * y; G[9][1] = -SHC30 * (3 * xx - 3 * yy);
    G[10][0] = SHC31 * y * z; G[10][1] = SHC31 * x * z; G[10][2] = SHC31 * x * y;
    G[11][0] = 2 * SHC32 * x * y; G[11][1] = -SHC32 * (4 * zz - xx - 3 * yy); G[11][2] = -8 * SHC32 * y * z;
    G[12][0] = -6 * SHC33 * x * z; G[12][1] = -6 * SHC33 * y * z; G[12][2] = SHC33 * (6 * zz - 3 * xx - 3 * yy);
    G[13][0] = -SHC32 * (4 * zz - 3 * xx - yy); G[13][1] = 2 * SHC32 * x * y; G[13][2] = -8 * SHC32 * x * z;
    G[14][0] = 2 * SHC34 * x * z; G[14][1] = -2 * SHC34 * y * z; G[14][2] = SHC34 * (xx - yy);
    G[15][0] = -SHC30 * (3 * xx - 3 * yy); G[15][1] = 6 * SHC30 * x * y;
}

static uint64_t fnv(uint64_t h, uint64_t v) {
    for (int i = 0; i < 8; ++i) { h ^= (v >> (8 * i)) & 0xFF; h *= 1099511628211ull; }
    return h;
}

static vis64 *g_sort_vis;
static int cmp_order(const void *pa, const void *pb) {
    int64_t a = *(const int64_t *)pa, b = *(const int64_t *)pb;
    uint64_t ka = g_sort_vis[a].key, kb = g_sort_vis[b].key;
    if (ka != kb) return ka < kb ? -1 : 1;
    return a < b ? -1 : (a > b);
}

static void state_free(scene_state *st) {
    free(st->vis); free(st->order); free(st->tstart); free(st->tlist);
}

/* Geometry of every Gaussian, culling, sorting and tile lists. */
static void prepare(const orc_scene64 *sc, const orc_camera *cam, const orc_config *cfg,
                    scene_state *st, uint64_t *sig) {
    const int64_t N = sc->n, W = cam->width, H = cam->height;
    const int64_t tw = cfg->tile_width, th = cfg->tile_height;
    const int64_t tx_n = (W + tw - 1) / tw, ty_n = (H + th - 1) / th;
    memset(st, 0, sizeof(*st));
    st->W = W; st->H = H; st->tiles_x = tx_n; st->tiles_y = ty_n;
    st->vis = (vis64 *)calloc((size_t)(N > 0 ? N : 1), sizeof(vis64));
    double qc[4] = {cam->rot[0], cam->rot[1], cam->rot[2], cam->rot[3]};
    double qn = sqrt(qc[0] * qc[0] + qc[1] * qc[1] + qc[2] * qc[2] + qc[3] * qc[3]);
    for (int k = 0; k < 4; ++k) qc[k] /= qn;
    double qi[4] = {qc[0], -qc[1], -qc[2], -qc[3]};
    mat_from_quat(qi, st->W9);
    const double *Wm = st->W9;
    const int ncoef = cfg->max_sh_degree == 0 ? 1 : cfg->max_sh_degree == 1 ? 4 :
                      cfg->max_sh_degree == 2 ? 9 : 16;
    int64_t V = 0;
    for (int64_t i = 0; i < N; ++i) {
        const double *P = sc->planes;
        double p[3] = {P[0 * N + i], P[1 * N + i], P[2 * N + i]};
        double s[3] = {P[3 * N + i], P[4 * N + i], P[5 * N + i]};
        double q[4] = {P[6 * N + i], P[7 * N + i], P[8 * N + i], P[9 * N + i]};
        double op = P[10 * N + i];
        if (!(op >= MIN_ALPHA)) continue;
        double rel[3] = {p[0] - cam->pos[0], p[1] - cam->pos[1], p[2] - cam->pos[2]};
        double pc[3];
        for (int r = 0; r < 3; ++r) pc[r] = Wm[3 * r] * rel[0] + Wm[3 * r + 1] * rel[1] + Wm[3 * r + 2] * rel[2];
        if (!(pc[2] >= cam->near_plane && pc[2] <= cam->far_plane)) continue;
        vis64 v;
        memset(&v, 0, sizeof(v));
        double z = pc[2];
        v.mx = cam->focal_x * pc[0] / z + cam->principal_x;
        v.my = cam->focal_y * pc[1] / z + cam->principal_y;
        double qnorm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        double qh[4] = {1, 0, 0, 0};
        if (qnorm > 1e-6) for (int k = 0; k < 4; ++k) qh[k] = q[k] / qnorm;
        v.qnorm = qnorm;
        memcpy(v.qhat, qh, sizeof(qh));
        mat_from_quat(qh, v.R);
        for (int k = 0; k < 3; ++k) {
            v.sclamp[k] = s[k] < 0.00001;
            v.s[k] = v.sclamp[k] ? 0.00001 : s[k];
        }
        double M[9];
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) M[3 * r + c] = v.R[3 * r + c] * v.s[c];
        double S[3][3];
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c)
            S[r][c] = M[3 * r] * M[3 * c] + M[3 * r + 1] * M[3 * c + 1] + M[3 * r + 2] * M[3 * c + 2];
        v.S[0] = S[0][0]; v.S[1] = S[0][1]; v.S[2] = S[0][2]; v.S[3] = S[1][1]; v.S[4] = S[1][2]; v.S[5] = S[2][2];
        double j00 = cam->focal_x / z, j02 = -cam->focal_x * pc[0] / (z * z);
        double j11 = cam->focal_y / z, j12 = -cam->focal_y * pc[1] / (z * z);
        for (int k = 0; k < 3; ++k) {
            v.t0[k] = j00 * Wm[k] + j02 * Wm[6 + k];
            v.t1[k] = j11 * Wm[3 + k] + j12 * Wm[6 + k];
        }
        double St0[3], St1[3];
        for (int r = 0; r < 3; ++r) {
            St0[r] = S[r][0] * v.t0[0] + S[r][1] * v.t0[1] + S[r][2] * v.t0[2];
            St1[r] = S[r][0] * v.t1[0] + S[r][1] * v.t1[1] + S[r][2] * v.t1[2];
        }
        double c00 = v.t0[0] * St0[0] + v.t0[1] * St0[1] + v.t0[2] * St0[2] + cfg->lowpass_filter;
        double c01 = v.t0[0] * St1[0] + v.t0[1] * St1[1] + v.t0[2] * St1[2];
        double c11 = v.t1[0] * St1[0] + v.t1[1] * St1[1] + v.t1[2] * St1[2] + cfg->lowpass_filter;
        v.c2[0] = c00; v.c2[1] = c01; v.c2[2] = c11;
        double mid = 0.5 * (c00 + c11), diff = 0.5 * (c00 - c11);
        double lmax = mid + sqrt(diff * diff + c01 * c01);
        double rf = ceil(3.0 * sqrt(lmax));
        if (!(rf >= 1.0)) continue;
        int64_t r = (int64_t)rf;
        if (v.mx + r < 0 || v.mx - r >= W || v.my + r < 0 || v.my - r >= H) continue;
        int64_t min_px = (int64_t)trunc(v.mx - r), max_px = (int64_t)trunc(v.mx + r);
        int64_t min_py = (int64_t)trunc(v.my - r), max_py = (int64_t)trunc(v.my + r);
        if (max_px < 0 || min_px >= W || max_py < 0 || min_py >= H) continue;
        if (min_px < 0) min_px = 0;
        if (max_px >= W) max_px = W - 1;
        if (min_py < 0) min_py = 0;
        if (max_py >= H) max_py = H - 1;
        v.rect[0] = min_px / tw; v.rect[1] = min_py / th; v.rect[2] = max_px / tw; v.rect[3] = max_py / th;
        double det = c00 * c11 - c01 * c01;
        v.det = det;
        if (det > -1e-7 && det < 1e-7) { v.inv_zero = 1; v.a = v.b = v.c = 0; }
        else { v.a = c11 / det; v.b = -c01 / det; v.c = c00 / det; }
        /* colour */
        double dl = sqrt(rel[0] * rel[0] + rel[1] * rel[1] + rel[2] * rel[2]);
        v.dirlen = dl;
        for (int k = 0; k < 3; ++k) v.dir[k] = dl > 1e-6 ? rel[k] / dl : 0.0;
        sh_basis64(v.dir, v.Y);
        const double *sh = sc->sh + i * 48;
        for (int ch = 0; ch < 3; ++ch) {
            double acc = 0.5;
            for (int k = 0; k < ncoef; ++k) acc += sh[3 * k + ch] * v.Y[k];
            v.colraw[ch] = acc;
            v.col[ch] = acc < 0 ? 0 : acc > 1 ? 1 : acc;
        }
        v.op = op;
        v.z = z;
        memcpy(v.pc, pc, sizeof(pc));
        double scaled = trunc(z * 1000.0);
        if (scaled < 0) scaled = 0;
        if (scaled > 2147483647.0) scaled = 2147483647.0;
        v.key = (uint64_t)scaled;
        v.orig = i;
        st->vis[V++] = v;
        if (sig) {
            *sig = fnv(*sig, (uint64_t)i);
            *sig = fnv(*sig, (uint64_t)r);
            for (int k = 0; k < 4; ++k) *sig = fnv(*sig, (uint64_t)v.rect[k]);
            for (int ch = 0; ch < 3; ++ch) *sig = fnv(*sig, (uint64_t)((v.colraw[ch] < 0) + 2 * (v.colraw[ch] > 1)));
            *sig = fnv(*sig, v.key);
        }
    }
    st->V = V;
    st->order = (int64_t *)malloc((size_t)(V > 0 ? V : 1) * sizeof(int64_t));
    for (int64_t k = 0; k < V; ++k) st->order[k] = k;
    g_sort_vis = st->vis;
    qsort(st->order, (size_t)V, sizeof(int64_t), cmp_order);
    const int64_t T = tx_n * ty_n;
    st->tstart = (int64_t *)calloc((size_t)(T + 1), sizeof(int64_t));
    for (int64_t k = 0; k < V; ++k) {
        vis64 *v = &st->vis[k];
        for (int64_t ty = v->rect[1]; ty <= v->rect[3]; ++ty)
            for (int64_t tx = v->rect[0]; tx <= v->rect[2]; ++tx) st->tstart[ty * tx_n + tx + 1]++;
    }
    for (int64_t t = 0; t < T; ++t) st->tstart[t + 1] += st->tstart[t];
    st->tlist = (int64_t *)malloc((size_t)(st->tstart[T] > 0 ? st->tstart[T] : 1) * sizeof(int64_t));
    int64_t *fillp = (int64_t *)malloc((size_t)T * sizeof(int64_t));
    for (int64_t t = 0; t < T; ++t) fillp[t] = st->tstart[t];
    for (int64_t k = 0; k < V; ++k) {
        int64_t vi = st->order[k];
        vis64 *v = &st->vis[vi];
        for (int64_t ty = v->rect[1]; ty <= v->rect[3]; ++ty)
            for (int64_t tx = v->rect[0]; tx <= v->rect[2]; ++tx) st->tlist[fillp[ty * tx_n + tx]++] = vi;
    }
    free(fillp);
}

typedef struct { int64_t vi; double alpha, G, T, power; int aclamp, pclamp; } contrib;

/* Forward blend of one pixel: the contributors in order; returns their count. */
static int64_t pixel_contribs(const scene_state *st, int64_t px, int64_t py, const orc_config *cfg,
                              contrib *out, double *T_final, uint64_t *sig) {
    (void)cfg;
    const int64_t tile = (py / cfg->tile_height) * st->tiles_x + px / cfg->tile_width;
    double pxc = (double)px + 0.5, pyc = (double)py + 0.5;
    double T = 1.0;
    int64_t n = 0;
    for (int64_t k = st->tstart[tile]; k < st->tstart[tile + 1]; ++k) {
        const vis64 *v = &st->vis[st->tlist[k]];
        double dx = pxc - v->mx, dy = pyc - v->my;
        double power = 0.5 * (v->a * dx * dx + 2.0 * v->b * dx * dy + v->c * dy * dy);
        int pclamp = 0;
        if (power < 0) { power = 0; pclamp = 1; }
        else if (power > 4.5) { if (sig) *sig = fnv(*sig, 1); continue; }
        double G = exp(-power);
        double alpha = v->op * G;
        int aclamp = 0;
        if (alpha > 0.99) { alpha = 0.99; aclamp = 1; }
        if (alpha < MIN_ALPHA) { if (sig) *sig = fnv(*sig, 2); continue; }
        if (sig) *sig = fnv(*sig, 3 + (uint64_t)aclamp + 4 * (uint64_t)pclamp);
        contrib c = {st->tlist[k], alpha, G, T, power, aclamp, pclamp};
        out[n++] = c;
        T *= (1.0 - alpha);
        if (T <= T_EPS) break;
    }
    *T_final = T;
    return n;
}

static int64_t max_list(const scene_state *st) {
    int64_t m = 1;
    const int64_t T = st->tiles_x * st->tiles_y;
    for (int64_t t = 0; t < T; ++t)
        if (st->tstart[t + 1] - st->tstart[t] > m) m = st->tstart[t + 1] - st->tstart[t];
    return m;
}

int orc64_forward(const orc_scene64 *sc, const orc_camera *cam, const orc_config *cfg,
                  double *rgb, double *depth, uint64_t *signature) {
    scene_state st;
    uint64_t sig = 1469598103934665603ull;
    prepare(sc, cam, cfg, &st, &sig);
    contrib *buf = (contrib *)malloc((size_t)max_list(&st) * sizeof(contrib));
    for (int64_t py = 0; py < st.H; ++py)
        for (int64_t px = 0; px < st.W; ++px) {
            double T;
            int64_t n = st.V ? pixel_contribs(&st, px, py, cfg, buf, &T, &sig) : 0;
            if (!st.V) T = 1.0;
            double C[3] = {0, 0, 0}, D = 0;
            for (int64_t k = 0; k < n; ++k) {
                const vis64 *v = &st.vis[buf[k].vi];
                double w = buf[k].alpha * buf[k].T;
                for (int ch = 0; ch < 3; ++ch) C[ch] += v->col[ch] * w;
                D += v->z * w;
            }
            if (cfg->white_background && st.V) for (int ch = 0; ch < 3; ++ch) C[ch] += T;
            for (int ch = 0; ch < 3; ++ch) rgb[(py * st.W + px) * 3 + ch] = C[ch];
            depth[py * st.W + px] = cfg->compute_depth ? D : 0.0;
        }
    free(buf);
    state_free(&st);
    if (signature) *signature = sig;
    return 0;
}

int orc64_backward(const orc_scene64 *sc, const orc_camera *cam, const orc_config *cfg,
                   const double *g_rgb, const double *g_depth, orc_grads64 *out,
                   double *g_vis /* optional V x 10: mx my a b c op r g b z */) {
    scene_state st;
    prepare(sc, cam, cfg, &st, NULL);
    const int64_t V = st.V, N = sc->n;
    double *gv = (double *)calloc((size_t)(V > 0 ? V : 1) * 10, sizeof(double));
    contrib *buf = (contrib *)malloc((size_t)max_list(&st) * sizeof(contrib));
    const int white = cfg->white_background != 0;
    for (int64_t py = 0; py < st.H && V > 0; ++py)
        for (int64_t px = 0; px < st.W; ++px) {
            double Tf;
            int64_t n = pixel_contribs(&st, px, py, cfg, buf, &Tf, NULL);
            const double *gC = g_rgb + (py * st.W + px) * 3;
            const double gD = cfg->compute_depth ? g_depth[py * st.W + px] : 0.0;
            double SC[3] = {0, 0, 0}, SD = 0;
            double pxc = (double)px + 0.5, pyc = (double)py + 0.5;
            for (int64_t k = n - 1; k >= 0; --k) {
                const contrib *c = &buf[k];
                const vis64 *v = &st.vis[c->vi];
                double *g = gv + c->vi * 10;
                double a = c->alpha, Tj = c->T, w = a * Tj;
                double dalpha = 0;
                for (int ch = 0; ch < 3; ++ch) {
                    g[6 + ch] += gC[ch] * w;
                    dalpha += gC[ch] * (v->col[ch] * Tj - SC[ch] / (1 - a));
                }
                g[9] += gD * w;
                dalpha += gD * (v->z * Tj - SD / (1 - a));
                if (white) dalpha -= (gC[0] + gC[1] + gC[2]) * Tf / (1 - a);
                for (int ch = 0; ch < 3; ++ch) SC[ch] += v->col[ch] * w;
                SD += v->z * w;
                if (c->aclamp) continue;
                g[5] += dalpha * c->G;
                double dpower = -dalpha * v->op * c->G;
                if (c->pclamp) continue;
                double dx = pxc - v->mx, dy = pyc - v->my;
                g[2] += dpower * 0.5 * dx * dx;
                g[3] += dpower * dx * dy;
                g[4] += dpower * 0.5 * dy * dy;
                g[0] += dpower * (-(v->a * dx + v->b * dy));
                g[1] += dpower * (-(v->b * dx + v->c * dy));
            }
        }
    free(buf);
    if (g_vis) memcpy(g_vis, gv, (size_t)V * 10 * sizeof(double));
    const int ncoef = cfg->max_sh_degree == 0 ? 1 : cfg->max_sh_degree == 1 ? 4 :
                      cfg->max_sh_degree == 2 ? 9 : 16;
    const double *Wm = st.W9;
    const double fx = cam->focal_x, fy = cam->focal_y;
    for (int64_t k = 0; k < V; ++k) {
        const vis64 *v = &st.vis[k];
        const double *g = gv + k * 10;
        const int64_t i = v->orig;
        double gpos[3] = {0, 0, 0}, gpc[3] = {0, 0, 0};
        out->planes[10 * N + i] += g[5];
        /* colour -> SH and view direction */
        const double *sh = sc->sh + i * 48;
        double gcol[3];
        for (int ch = 0; ch < 3; ++ch) gcol[ch] = (v->colraw[ch] < 0 || v->colraw[ch] > 1) ? 0 : g[6 + ch];
        double dY[16];
        for (int kk = 0; kk < 16; ++kk) {
            dY[kk] = 0;
            if (kk >= ncoef) continue;
            for (int ch = 0; ch < 3; ++ch) {
                out->sh[i * 48 + 3 * kk + ch] += gcol[ch] * v->Y[kk];
                dY[kk] += gcol[ch] * sh[3 * kk + ch];
            }
        }
        if (v->dirlen > 1e-6) {
            double GY[16][3], gd[3] = {0, 0, 0};
            sh_basis_grad64(v->dir, GY);
            for (int kk = 0; kk < ncoef; ++kk) for (int a = 0; a < 3; ++a) gd[a] += dY[kk] * GY[kk][a];
            double dot = gd[0] * v->dir[0] + gd[1] * v->dir[1] + gd[2] * v->dir[2];
            for (int a = 0; a < 3; ++a) gpos[a] += (gd[a] - v->dir[a] * dot) / v->dirlen;
        }
        /* depth and screen position */
        const double x = v->pc[0], y = v->pc[1], z = v->pc[2];
        gpc[2] += g[9];
        gpc[0] += g[0] * fx / z;
        gpc[1] += g[1] * fy / z;
        gpc[2] += -g[0] * fx * x / (z * z) - g[1] * fy * y / (z * z);
        out->viewspace[i] += sqrt(g[0] * g[0] + g[1] * g[1]);
        out->count[i] += 1;
        /* inverse covariance -> screen covariance */
        double gc00 = 0, gc01 = 0, gc11 = 0;
        if (!v->inv_zero) {
            const double c00 = v->c2[0], c01 = v->c2[1], c11 = v->c2[2], D = v->det;
            const double K = (-g[2] * c11 + g[3] * c01 - g[4] * c00) / (D * D);
            gc00 = g[4] / D + K * c11;
            gc11 = g[2] / D + K * c00;
            gc01 = -g[3] / D - 2.0 * K * c01;
        }
        /* screen covariance -> T rows and world covariance */
        double S[3][3] = {{v->S[0], v->S[1], v->S[2]}, {v->S[1], v->S[3], v->S[4]}, {v->S[2], v->S[4], v->S[5]}};
        double St0[3], St1[3];
        for (int r = 0; r < 3; ++r) {
            St0[r] = S[r][0] * v->t0[0] + S[r][1] * v->t0[1] + S[r][2] * v->t0[2];
            St1[r] = S[r][0] * v->t1[0] + S[r][1] * v->t1[1] + S[r][2] * v->t1[2];
        }
        double gt0[3], gt1[3], F[3][3];
        for (int r = 0; r < 3; ++r) {
            gt0[r] = 2 * gc00 * St0[r] + gc01 * St1[r];
            gt1[r] = gc01 * St0[r] + 2 * gc11 * St1[r];
        }
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c)
            F[r][c] = gc00 * v->t0[r] * v->t0[c] + 0.5 * gc01 * (v->t0[r] * v->t1[c] + v->t1[r] * v->t0[c]) +
                      gc11 * v->t1[r] * v->t1[c];
        /* T = J W -> J -> camera-space position */
        double gj00 = gt0[0] * Wm[0] + gt0[1] * Wm[1] + gt0[2] * Wm[2];
        double gj02 = gt0[0] * Wm[6] + gt0[1] * Wm[7] + gt0[2] * Wm[8];
        double gj11 = gt1[0] * Wm[3] + gt1[1] * Wm[4] + gt1[2] * Wm[5];
        double gj12 = gt1[0] * Wm[6] + gt1[1] * Wm[7] + gt1[2] * Wm[8];
        const double z2 = z * z, z3 = z2 * z;
        gpc[2] += gj00 * (-fx / z2) + gj02 * (2 * fx * x / z3) + gj11 * (-fy / z2) + gj12 * (2 * fy * y / z3);
        gpc[0] += gj02 * (-fx / z2);
        gpc[1] += gj12 * (-fy / z2);
        for (int a = 0; a < 3; ++a) gpos[a] += Wm[a] * gpc[0] + Wm[3 + a] * gpc[1] + Wm[6 + a] * gpc[2];
        for (int a = 0; a < 3; ++a) out->planes[a * N + i] += gpos[a];
        /* world covariance -> scale and rotation: Sigma = M M^T, M = R diag(s) */
        double M[9], gM[9], gR[9];
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) M[3 * r + c] = v->R[3 * r + c] * v->s[c];
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c)
            gM[3 * r + c] = 2 * (F[r][0] * M[c] + F[r][1] * M[3 + c] + F[r][2] * M[6 + c]);
        for (int c = 0; c < 3; ++c) {
            double gs = gM[c] * v->R[c] + gM[3 + c] * v->R[3 + c] + gM[6 + c] * v->R[6 + c];
            if (!v->sclamp[c]) out->planes[(3 + c) * N + i] += gs;
        }
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) gR[3 * r + c] = gM[3 * r + c] * v->s[c];
        const double w = v->qhat[0], qx = v->qhat[1], qy = v->qhat[2], qz = v->qhat[3];
        double gq[4];
        gq[0] = 2 * (-qz * gR[1] + qy * gR[2] + qz * gR[3] - qx * gR[5] - qy * gR[6] + qx * gR[7]);
        gq[1] = 2 * (qy * gR[1] + qz * gR[2] + qy * gR[3] - 2 * qx * gR[4] - w * gR[5] + qz * gR[6] + w * gR[7] - 2 * qx * gR[8]);
        gq[2] = 2 * (-2 * qy * gR[0] + qx * gR[1] + w * gR[2] + qx * gR[3] + qz * gR[5] - w * gR[6] + qz * gR[7] - 2 * qy * gR[8]);
        gq[3] = 2 * (-2 * qz * gR[0] - w * gR[1] + qx * gR[2] + w * gR[3] - 2 * qz * gR[4] + qy * gR[5] + qx * gR[6] + qy * gR[7]);
        if (v->qnorm > 1e-6) {
            double dot = gq[0] * w + gq[1] * qx + gq[2] * qy + gq[3] * qz;
            for (int a = 0; a < 4; ++a) out->planes[(6 + a) * N + i] += (gq[a] - v->qhat[a] * dot) / v->qnorm;
        }
    }
    free(gv);
    int64_t Vout = V;
    state_free(&st);
    return (int)Vout;
}
