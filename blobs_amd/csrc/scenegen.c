/*
 * scenegen.c -- deterministic synthetic scenes for the benchmark and tests.
 *
 * RNG: Marsaglia xorshift128, state seeded like rand_xorshift's
 * XorShiftRng::from_seed([k; 16]) (the reference's only seeding convention,
 * blobs/benches/spatial_hash.rs:9): x = y = z = w = k * 0x01010101.
 *   U[a,b]  = a + (b - a) * u,  u = (next >> 8) * 2^-24     (f64, then f32)
 *   N(0,s)  = Box-Muller in f64 (two uniforms), then f32
 *   logU    = exp(U[ln a, ln b])
 * Overlap rejection redraws a centre until it overlaps nothing already
 * placed (uniform grid); after max_tries the last draw is kept and counted.
 * The RNG state is passed in and out (uint32_t[4]) so callers can chain.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static uint32_t xs_next(uint32_t st[4]) {
    uint32_t x = st[0];
    uint32_t t = x ^ (x << 11);
    st[0] = st[1];
    st[1] = st[2];
    st[2] = st[3];
    uint32_t w = st[3];
    st[3] = w ^ (w >> 19) ^ (t ^ (t >> 8));
    return st[3];
}

static double xs_u01(uint32_t st[4]) { return (double)(xs_next(st) >> 8) * (1.0 / 16777216.0); }
static double xs_uniform(uint32_t st[4], double a, double b) { return a + (b - a) * xs_u01(st); }
static double xs_normal(uint32_t st[4], double sigma) {
    double u1 = xs_u01(st), u2 = xs_u01(st);
    return sigma * sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
}

void sg_seed(uint32_t st[4], uint32_t k) {
    uint32_t v = (k & 0xffu) * 0x01010101u;
    st[0] = st[1] = st[2] = st[3] = v;
    if (v == 0) st[0] = 0x0BAD5EEDu; /* xorshift must not start at all-zero */
}

/* n draws of U[a,b] (log_scale: exp(U[ln a, ln b])) into out (f32). */
void sg_uniform(uint32_t st[4], int64_t n, double a, double b, int log_scale, float *out) {
    for (int64_t i = 0; i < n; ++i) {
        if (log_scale)
            out[i] = (float)exp(xs_uniform(st, log(a), log(b)));
        else
            out[i] = (float)xs_uniform(st, a, b);
    }
}

/* n draws of N(0, sigma) into out (f32). */
void sg_normal(uint32_t st[4], int64_t n, double sigma, float *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = (float)xs_normal(st, sigma);
}

/* uniform integers in [lo, hi] */
void sg_randint(uint32_t st[4], int64_t n, int64_t lo, int64_t hi, int32_t *out) {
    for (int64_t i = 0; i < n; ++i) {
        int64_t span = hi - lo + 1;
        int64_t v = lo + (int64_t)(xs_u01(st) * (double)span);
        if (v > hi) v = hi;
        out[i] = (int32_t)v;
    }
}

/* ---- occupancy grid over discs -------------------------------------- */
typedef struct {
    double x0, y0, inv;
    int32_t nx, ny;
    int32_t *head, *next;
    const float *cx, *cy, *r;
    int64_t cap;
} occ_t;

static void occ_init(occ_t *g, double x0, double y0, double x1, double y1, double cell, int64_t cap) {
    g->x0 = x0;
    g->y0 = y0;
    g->inv = 1.0 / cell;
    double w = x1 - x0, h = y1 - y0;
    g->nx = (int32_t)(w / cell) + 1;
    g->ny = (int32_t)(h / cell) + 1;
    while ((double)g->nx * g->ny > 4.0 * (double)cap + 4096.0) {
        cell *= 1.5;
        g->inv = 1.0 / cell;
        g->nx = (int32_t)(w / cell) + 1;
        g->ny = (int32_t)(h / cell) + 1;
    }
    g->head = (int32_t *)malloc(sizeof(int32_t) * (size_t)g->nx * (size_t)g->ny);
    for (int64_t i = 0; i < (int64_t)g->nx * g->ny; ++i) g->head[i] = -1;
    g->next = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
    g->cap = cap;
}

static void occ_free(occ_t *g) {
    free(g->head);
    free(g->next);
}

static int32_t occ_cx(occ_t *g, double x) {
    double f = floor((x - g->x0) * g->inv);
    if (f < 0) return 0;
    if (f > g->nx - 1) return g->nx - 1;
    return (int32_t)f;
}
static int32_t occ_cy(occ_t *g, double y) {
    double f = floor((y - g->y0) * g->inv);
    if (f < 0) return 0;
    if (f > g->ny - 1) return g->ny - 1;
    return (int32_t)f;
}

static void occ_add(occ_t *g, int32_t id, double x, double y) {
    int32_t c = occ_cy(g, y) * g->nx + occ_cx(g, x);
    g->next[id] = g->head[c];
    g->head[c] = id;
}

/* does disc (x, y, r) overlap any stored disc (owner != skip_owner)? search radius reach */
static int occ_hits(occ_t *g, double x, double y, double r, double reach, const int32_t *owner, int32_t skip) {
    int32_t cx0 = occ_cx(g, x - reach), cx1 = occ_cx(g, x + reach);
    int32_t cy0 = occ_cy(g, y - reach), cy1 = occ_cy(g, y + reach);
    for (int32_t cy = cy0; cy <= cy1; ++cy)
        for (int32_t cx = cx0; cx <= cx1; ++cx)
            for (int32_t k = g->head[cy * g->nx + cx]; k >= 0; k = g->next[k]) {
                if (owner && owner[k] == skip) continue;
                double dx = x - g->cx[k], dy = y - g->cy[k], rr = r + g->r[k];
                if (dx * dx + dy * dy < rr * rr) return 1;
            }
    return 0;
}

/*
 * n single-sphere bodies: for each, radius ~ U[rmin, rmax], then centre ~ U
 * over the box (redrawn while it overlaps, if reject), then velocity
 * N(0, vel_sigma) per component.  Returns the number of spheres whose
 * placement gave up after max_tries.
 */
int64_t sg_spheres(uint32_t st[4], int64_t n, double rmin, double rmax, double bx0, double by0, double bx1,
                   double by1, double vel_sigma, int reject, int64_t max_tries, float *pos, float *vel,
                   float *radius) {
    float *cx = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    float *cy = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    occ_t g;
    occ_init(&g, bx0, by0, bx1, by1, 2.0 * rmax, n);
    g.cx = cx;
    g.cy = cy;
    g.r = radius;
    int64_t forced = 0;
    for (int64_t i = 0; i < n; ++i) {
        radius[i] = (float)xs_uniform(st, rmin, rmax);
        double x = 0, y = 0;
        int64_t t = 0;
        for (;;) {
            x = (double)(float)xs_uniform(st, bx0, bx1);
            y = (double)(float)xs_uniform(st, by0, by1);
            if (!reject || !occ_hits(&g, x, y, radius[i], radius[i] + rmax, NULL, -1)) break;
            if (++t >= max_tries) {
                ++forced;
                break;
            }
        }
        cx[i] = (float)x;
        cy[i] = (float)y;
        occ_add(&g, (int32_t)i, x, y);
        pos[2 * i] = (float)x;
        pos[2 * i + 1] = (float)y;
        vel[2 * i] = (float)xs_normal(st, vel_sigma);
        vel[2 * i + 1] = (float)xs_normal(st, vel_sigma);
    }
    occ_free(&g);
    free(cx);
    free(cy);
    return forced;
}

/*
 * Multi-collider bodies: count ~ U{cmin..cmax}; each collider radius ~
 * U[rmin, rmax] and offset uniform in a disc of radius off_r (r = R sqrt(u),
 * theta = 2 pi u'); body centre ~ U over the box, redrawn while any of its
 * colliders overlaps a collider of another body.  col_* arrays must hold
 * n * cmax entries; ncol[i] receives the count.  Returns total colliders.
 */
int64_t sg_multi(uint32_t st[4], int64_t n, int32_t cmin, int32_t cmax, double off_r, double rmin, double rmax,
                 double bx0, double by0, double bx1, double by1, double vel_sigma, int64_t max_tries, float *pos,
                 float *vel, int32_t *ncol, float *col_off, float *col_r, int64_t *forced_out) {
    int64_t cap = n * (int64_t)cmax;
    float *ax = (float *)malloc(sizeof(float) * (size_t)(cap > 0 ? cap : 1));
    float *ay = (float *)malloc(sizeof(float) * (size_t)(cap > 0 ? cap : 1));
    int32_t *owner = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
    occ_t g;
    occ_init(&g, bx0, by0, bx1, by1, 2.0 * rmax, cap);
    g.cx = ax;
    g.cy = ay;
    g.r = col_r;
    int64_t nc = 0, forced = 0;
    for (int64_t i = 0; i < n; ++i) {
        int32_t k = cmin + (int32_t)(xs_u01(st) * (double)(cmax - cmin + 1));
        if (k > cmax) k = cmax;
        ncol[i] = k;
        for (int32_t q = 0; q < k; ++q) {
            col_r[nc + q] = (float)xs_uniform(st, rmin, rmax);
            double rr = off_r * sqrt(xs_u01(st));
            double th = 6.283185307179586 * xs_u01(st);
            col_off[2 * (nc + q)] = (float)(rr * cos(th));
            col_off[2 * (nc + q) + 1] = (float)(rr * sin(th));
        }
        double x = 0, y = 0;
        int64_t t = 0;
        for (;;) {
            x = (double)(float)xs_uniform(st, bx0, bx1);
            y = (double)(float)xs_uniform(st, by0, by1);
            int hit = 0;
            for (int32_t q = 0; q < k && !hit; ++q) {
                double px = (double)((float)x + col_off[2 * (nc + q)]);
                double py = (double)((float)y + col_off[2 * (nc + q) + 1]);
                hit = occ_hits(&g, px, py, col_r[nc + q], col_r[nc + q] + rmax, owner, (int32_t)i);
            }
            if (!hit) break;
            if (++t >= max_tries) {
                ++forced;
                break;
            }
        }
        pos[2 * i] = (float)x;
        pos[2 * i + 1] = (float)y;
        vel[2 * i] = (float)xs_normal(st, vel_sigma);
        vel[2 * i + 1] = (float)xs_normal(st, vel_sigma);
        for (int32_t q = 0; q < k; ++q) {
            int64_t c = nc + q;
            ax[c] = (float)x + col_off[2 * c];
            ay[c] = (float)y + col_off[2 * c + 1];
            owner[c] = (int32_t)i;
            occ_add(&g, (int32_t)c, ax[c], ay[c]);
        }
        nc += k;
    }
    occ_free(&g);
    free(ax);
    free(ay);
    free(owner);
    if (forced_out) *forced_out = forced;
    return nc;
}

/*
 * m random pairs (a, b), a != b, with |pos_a - pos_b| < max_dist: a uniform
 * over bodies, b uniform among a's neighbours within max_dist (a redrawn if
 * it has none).  rest[k] = |pos_b - pos_a| in f32 (glam length).
 */
int64_t sg_pairs(uint32_t st[4], int64_t n, const float *pos, int64_t m, double max_dist, int32_t *pa, int32_t *pb,
                 float *rest) {
    if (n < 2) return 0;
    double minx = pos[0], maxx = pos[0], miny = pos[1], maxy = pos[1];
    for (int64_t i = 1; i < n; ++i) {
        if (pos[2 * i] < minx) minx = pos[2 * i];
        if (pos[2 * i] > maxx) maxx = pos[2 * i];
        if (pos[2 * i + 1] < miny) miny = pos[2 * i + 1];
        if (pos[2 * i + 1] > maxy) maxy = pos[2 * i + 1];
    }
    occ_t g;
    occ_init(&g, minx, miny, maxx, maxy, max_dist, n);
    for (int64_t i = 0; i < n; ++i) occ_add(&g, (int32_t)i, pos[2 * i], pos[2 * i + 1]);
    int32_t *cand = (int32_t *)malloc(sizeof(int32_t) * 4096);
    int64_t cand_cap = 4096, made = 0, guard = 0;
    while (made < m && guard < 1000 * (m + 1)) {
        ++guard;
        int64_t a = (int64_t)(xs_u01(st) * (double)n);
        if (a >= n) a = n - 1;
        double x = pos[2 * a], y = pos[2 * a + 1];
        int64_t nc = 0;
        int32_t cx0 = occ_cx(&g, x - max_dist), cx1 = occ_cx(&g, x + max_dist);
        int32_t cy0 = occ_cy(&g, y - max_dist), cy1 = occ_cy(&g, y + max_dist);
        for (int32_t cy = cy0; cy <= cy1; ++cy)
            for (int32_t cx = cx0; cx <= cx1; ++cx)
                for (int32_t k = g.head[cy * g.nx + cx]; k >= 0; k = g.next[k]) {
                    if (k == a) continue;
                    double dx = pos[2 * k] - x, dy = pos[2 * k + 1] - y;
                    if (dx * dx + dy * dy >= max_dist * max_dist) continue;
                    if (nc == cand_cap) {
                        cand_cap *= 2;
                        cand = (int32_t *)realloc(cand, sizeof(int32_t) * (size_t)cand_cap);
                    }
                    cand[nc++] = k;
                }
        if (nc == 0) continue;
        /* deterministic order regardless of grid layout */
        for (int64_t i = 1; i < nc; ++i) {
            int32_t v = cand[i];
            int64_t j = i - 1;
            while (j >= 0 && cand[j] > v) {
                cand[j + 1] = cand[j];
                --j;
            }
            cand[j + 1] = v;
        }
        int64_t pick = (int64_t)(xs_u01(st) * (double)nc);
        if (pick >= nc) pick = nc - 1;
        int32_t b = cand[pick];
        pa[made] = (int32_t)a;
        pb[made] = b;
        float dx = pos[2 * b] - pos[2 * a], dy = pos[2 * b + 1] - pos[2 * a + 1];
        rest[made] = sqrtf(dx * dx + dy * dy);
        ++made;
    }
    free(cand);
    occ_free(&g);
    return made;
}
