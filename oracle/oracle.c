/*
 * oracle.c -- CPU restatement of ResearchUV's LSCM unfold path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the MI355X engine
 * under researchuv_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  The product never links, imports or
 * calls it, and nothing here runs on a GPU.
 *
 * Each function restates the reference algorithm step by step (same operation
 * order, same tie-breaks, same constants) and cites the reference file:line it
 * follows.  Compile with -ffp-contract=off so no multiply-add is fused: the
 * reference (Rust) never contracts.
 *
 * Pinning status (what checks each part; the tests say the same):
 *   orc_cube / orc_grid ............ counts 294/432, 11*9/160 (meshgen.rs:330-368)
 *   orc_weld ....................... 218/432 cube (parity.rs:31-33), seam + degenerate tests (weld.rs:69-125)
 *   orc_segment .................... 6x49 cube charts, flat grid / triangle / tetra loops (segment.rs:456-510),
 *                                    cylinder 2 loops (meshgen.rs:370-380), annulus border 128 (parity.rs:155),
 *                                    hand-walked loops through a pinch vertex: straightest turn, strict-max
 *                                    tie keeping the smaller id, smallest first step (segment.rs:258-330)
 *   orc_eig3 / orc_tangent_frame ... known spectra (lscm_eig.rs:72-120), numpy.linalg.eigh, planar-isometry check,
 *                                    one-rotation sign convention in each plane, a hand-derived two-rotation
 *                                    case fixing the accumulation order V' = Q V (lscm_eig.rs:57-62)
 *   orc_chart_system (pins, weights, assembly) ... cube closed form (L = 72 x 5-point Laplacian, pins on the
 *                                    unit-square perimeter), symmetric/row-sum invariants, dense numpy rebuild
 *   orc_sparse_solve ............... sparse.rs:272-423 known answers, numpy.linalg.solve brute force
 *   orc_unfold_chart ............... cube affine closed form, iters == 2, dense-solve parity
 *   orc_rect_point / orc_anchor_ids  perimeter walk known answers, farthest-point anchors on a square
 *   orc_par_unfold ................. bitwise equal for 1, 2, 3 and 8 workers and to the serial driver
 *   orc_lscm_loss_frozen ........... affine maps on a right-triangle grid: uv == A x + t, closed-form energy
 *                                    2 nx ny (hy/hx |A e_x|^2 + hx/hy |A e_y|^2) plus the target term
 *   orc_lscm_vjp_frozen, orc_lscm_grad_scatter .. central differences of orc_lscm_loss_frozen; rigid-motion
 *                                    invariance (sum of gradients 0, sum p x grad 0); scale: sum p . grad = -2 E
 *   orc_cg_solve ................... 1-D Laplacian x == 1 (solver.rs:767-779)
 *   orc_ic0_* ...................... L*L^T == A on the pattern (precond.rs:708-737), level partition (:739-755),
 *                                    dense triangular solve of the same factor
 *   orc_coloring ................... distance-1 validity, colour bound (solver.rs:889-909)
 *   orc_winding_flips .............. recut.rs:155-169 known answers, a planted fold
 *   orc_chart_distortion ........... identity unfold of a triangle is conformal (metrics.rs:239-257); affine maps
 *                                    of a planar mesh (closed-form singular values and area ratio); degenerate
 *                                    triangles counted as flips
 *   orc_rectangularize ............. unit square on its border (metrics.rs:258-279); a rectangle rotated by a
 *                                    ladder angle comes back axis-aligned with its side lengths; border-free
 *                                    normalisation to [0,1]^2
 *   Every function above has a check independent of its restatement.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define EPS_DEGENERATE 1.1754943508222875e-38 /* lscm.rs:36 */
#define AREA_HALF 0.5                          /* lscm.rs:38 */
#define STAGNATION 0.01                        /* lscm.rs:44 */
#define DEFAULT_MAX_ITER 100                   /* lscm.rs:46 */
#define DEFAULT_MAX_ITER_GUARD 10              /* lscm.rs:48 */

/* ------------------------------------------------------------------ vec */
typedef struct { double x, y, z; } v3;
typedef struct { double u, v; } v2;

static v3 v3sub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static v3 v3add(v3 a, v3 b) { v3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static v3 v3div(v3 a, double s) { v3 r = {a.x / s, a.y / s, a.z / s}; return r; }
/* vec.rs:104-106: x*x + y*y + z*z, left to right */
static double v3dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* vec.rs:108-114 */
static v3 v3cross(v3 a, v3 o) {
    v3 r = {a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x};
    return r;
}
static double v3len(v3 a) { return sqrt(v3dot(a, a)); }
static v3 P3(const double* p, size_t i) { v3 r = {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; return r; }

/* dynamic arrays */
typedef struct { uint32_t* a; size_t n, cap; } vu32;
static void vu32_push(vu32* v, uint32_t x) {
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 8; v->a = (uint32_t*)realloc(v->a, v->cap * sizeof(uint32_t)); }
    v->a[v->n++] = x;
}

/* ------------------------------------------------------------------ meshgen */
/* meshgen.rs:12-54 */
int64_t orc_cube(int n, double* pos, uint32_t* faces) {
    static const int defs[6][4] = {{2, 0, 0, 1}, {2, 1, 0, 1}, {0, 0, 1, 2}, {0, 1, 1, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}};
    size_t np = 0, nf = 0;
    for (int d = 0; d < 6; ++d) {
        int fx = defs[d][0], va = defs[d][2], vb = defs[d][3];
        double fv = (double)defs[d][1];
        uint32_t base = (uint32_t)np;
        for (int i = 0; i <= n; ++i)
            for (int j = 0; j <= n; ++j) {
                double pt[3];
                for (int k = 0; k < 3; ++k) {
                    if (k == fx) pt[k] = fv;
                    else if (k == va) pt[k] = (double)i / (double)n;
                    else if (k == vb) pt[k] = (double)j / (double)n;
                    else pt[k] = 0.0;
                }
                if (pos) { pos[3 * np] = pt[0]; pos[3 * np + 1] = pt[1]; pos[3 * np + 2] = pt[2]; }
                np++;
            }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                uint32_t a = base + (uint32_t)i * (uint32_t)(n + 1) + (uint32_t)j, b = a + 1, c = a + (uint32_t)(n + 1), dd = c + 1;
                if (faces) {
                    uint32_t* f = faces + 3 * nf;
                    f[0] = a; f[1] = b; f[2] = dd; f[3] = a; f[4] = dd; f[5] = c;
                }
                nf += 2;
            }
    }
    return ((int64_t)np << 32) | (int64_t)nf;
}

/* meshgen.rs:135-155 */
void orc_grid(int w, int h, double* pos, uint32_t* faces) {
    size_t np = 0;
    for (int j = 0; j <= h; ++j)
        for (int i = 0; i <= w; ++i) {
            pos[3 * np] = (double)i; pos[3 * np + 1] = (double)j; pos[3 * np + 2] = 0.0; np++;
        }
    uint32_t stride = (uint32_t)(w + 1);
    size_t nf = 0;
    for (int j = 0; j < h; ++j)
        for (int i = 0; i < w; ++i) {
            uint32_t a = (uint32_t)j * stride + (uint32_t)i, b = a + 1, c = a + stride, d = c + 1;
            uint32_t* f = faces + 3 * nf;
            f[0] = a; f[1] = b; f[2] = d; f[3] = a; f[4] = d; f[5] = c;
            nf += 2;
        }
}

/* ------------------------------------------------------------------ weld */
typedef struct { int64_t k[3]; uint32_t i; } weldkey;
static int64_t sat_i64(double v) {
    if (v != v) return 0;
    if (v >= 9.2233720368547758e18) return INT64_MAX;
    if (v <= -9.2233720368547758e18) return INT64_MIN;
    return (int64_t)v;
}
static int cmp_weldkey(const void* pa, const void* pb) {
    const weldkey* a = (const weldkey*)pa; const weldkey* b = (const weldkey*)pb;
    for (int t = 0; t < 3; ++t) { if (a->k[t] < b->k[t]) return -1; if (a->k[t] > b->k[t]) return 1; }
    return a->i < b->i ? -1 : (a->i > b->i);
}
/* weld.rs:19-57: canonical id = order of first occurrence of the rounded key. Returns welded vertex count;
 * pos_out must hold nv*3, faces_out nf*3; *nf_out receives the kept face count. */
int orc_weld(const double* pos, int nv, const uint32_t* faces, int nf, double tol, double* pos_out,
             uint32_t* faces_out, int* nf_out) {
    if (tol <= 0.0) {
        memcpy(pos_out, pos, sizeof(double) * 3 * (size_t)nv);
        memcpy(faces_out, faces, sizeof(uint32_t) * 3 * (size_t)nf);
        *nf_out = nf;
        return nv;
    }
    double inv = 1.0 / tol;
    weldkey* ks = (weldkey*)malloc(sizeof(weldkey) * (size_t)(nv ? nv : 1));
    for (int i = 0; i < nv; ++i) {
        ks[i].k[0] = sat_i64(round(pos[3 * i] * inv));
        ks[i].k[1] = sat_i64(round(pos[3 * i + 1] * inv));
        ks[i].k[2] = sat_i64(round(pos[3 * i + 2] * inv));
        ks[i].i = (uint32_t)i;
    }
    qsort(ks, (size_t)nv, sizeof(weldkey), cmp_weldkey);
    uint32_t* first = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nv ? nv : 1)); /* vertex -> first index of group */
    for (int g = 0; g < nv;) {
        int e = g;
        while (e < nv && ks[e].k[0] == ks[g].k[0] && ks[e].k[1] == ks[g].k[1] && ks[e].k[2] == ks[g].k[2]) e++;
        for (int t = g; t < e; ++t) first[ks[t].i] = ks[g].i;
        g = e;
    }
    uint32_t* remap = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nv ? nv : 1));
    int nw = 0;
    for (int i = 0; i < nv; ++i) {
        if (first[i] == (uint32_t)i) {
            remap[i] = (uint32_t)nw;
            memcpy(pos_out + 3 * nw, pos + 3 * i, sizeof(double) * 3);
            nw++;
        } else {
            remap[i] = remap[first[i]];
        }
    }
    int kept = 0;
    for (int f = 0; f < nf; ++f) {
        uint32_t a = remap[faces[3 * f]], b = remap[faces[3 * f + 1]], c = remap[faces[3 * f + 2]];
        if (a != b && b != c && a != c) {
            faces_out[3 * kept] = a; faces_out[3 * kept + 1] = b; faces_out[3 * kept + 2] = c; kept++;
        }
    }
    *nf_out = kept;
    free(ks); free(first); free(remap);
    return nw;
}

/* ------------------------------------------------------------------ segment */
typedef struct { uint32_t a, b; uint32_t face; uint32_t ord; } efrec;
static int cmp_efrec(const void* pa, const void* pb) {
    const efrec* x = (const efrec*)pa; const efrec* y = (const efrec*)pb;
    if (x->a != y->a) return x->a < y->a ? -1 : 1;
    if (x->b != y->b) return x->b < y->b ? -1 : 1;
    return x->ord < y->ord ? -1 : (x->ord > y->ord);
}

typedef struct {
    int nfaces; uint32_t* face_ids;
    int nverts; uint32_t* vertex_ids;
    uint32_t* tris;
    int nborder; uint32_t* border;
    int nloops; int* loop_len; uint32_t** loops;
    double area3d;
} orc_chart;

typedef struct {
    int n; orc_chart* ch;
    int ncut; uint32_t* cut; /* 2*ncut sorted keys */
} orc_charts;

/* sorted unique edge table with per-edge face list (edge_faces, segment.rs:73-81) */
typedef struct {
    size_t ne;
    uint32_t* ka; uint32_t* kb;     /* keys */
    size_t* off;                     /* face list offsets (ne+1) */
    uint32_t* faces;                 /* faces in insertion order */
} edgetab;

static void edgetab_build(const uint32_t* faces, int nf, edgetab* t) {
    size_t m = 3 * (size_t)nf;
    efrec* r = (efrec*)malloc(sizeof(efrec) * (m ? m : 1));
    for (int f = 0; f < nf; ++f) {
        uint32_t v[3] = {faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]};
        for (int k = 0; k < 3; ++k) {
            uint32_t x = v[k], y = v[(k + 1) % 3];
            efrec* e = &r[3 * (size_t)f + k];
            e->a = x < y ? x : y; e->b = x < y ? y : x; e->face = (uint32_t)f; e->ord = (uint32_t)(3 * f + k);
        }
    }
    qsort(r, m, sizeof(efrec), cmp_efrec);
    t->ka = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
    t->kb = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
    t->off = (size_t*)malloc(sizeof(size_t) * (m + 1));
    t->faces = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
    size_t ne = 0;
    for (size_t i = 0; i < m; ++i) {
        if (i == 0 || r[i].a != r[i - 1].a || r[i].b != r[i - 1].b) {
            t->ka[ne] = r[i].a; t->kb[ne] = r[i].b; t->off[ne] = i; ne++;
        }
        t->faces[i] = r[i].face;
    }
    t->off[ne] = m;
    t->ne = ne;
    free(r);
}
static void edgetab_free(edgetab* t) { free(t->ka); free(t->kb); free(t->off); free(t->faces); }
static long edgetab_find(const edgetab* t, uint32_t a, uint32_t b) {
    size_t lo = 0, hi = t->ne;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (t->ka[mid] < a || (t->ka[mid] == a && t->kb[mid] < b)) lo = mid + 1; else hi = mid;
    }
    if (lo < t->ne && t->ka[lo] == a && t->kb[lo] == b) return (long)lo;
    return -1;
}

/* segment.rs:104-117 */
static v3 face_normal_unit(const double* pos, const uint32_t* faces, uint32_t fi) {
    v3 a = P3(pos, faces[3 * fi]), b = P3(pos, faces[3 * fi + 1]), c = P3(pos, faces[3 * fi + 2]);
    v3 n = v3cross(v3sub(a, b), v3sub(c, b)); /* vec.rs:100-102 cross(a-b, c-b) */
    double l = v3len(n);
    if (l < 1e-30) { v3 z = {0.0, 0.0, 1.0}; return z; }
    return v3div(n, l);
}
/* segment.rs:61-70 */
static double face_area3d(const double* pos, const uint32_t* faces, uint32_t f) {
    v3 a = P3(pos, faces[3 * f]), b = P3(pos, faces[3 * f + 1]), c = P3(pos, faces[3 * f + 2]);
    return 0.5 * v3len(v3cross(v3sub(b, a), v3sub(c, a)));
}

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b; return x < y ? -1 : (x > y);
}
static long bsearch_u32(const uint32_t* a, size_t n, uint32_t key) {
    size_t lo = 0, hi = n;
    while (lo < hi) { size_t m = (lo + hi) / 2; if (a[m] < key) lo = m + 1; else hi = m; }
    return (lo < n && a[lo] == key) ? (long)lo : -1;
}

/* segment.rs:227-331 keep-going-straight walk. is_bnd[e] marks cut-or-boundary edges of the table. */
static void chart_border_loops(const double* pos, const uint32_t* faces, const uint32_t* comp, int ncomp,
                               const edgetab* et, const uint8_t* is_bnd, orc_chart* ch) {
    /* bnd_edges: sorted unique edge indices (edge table is sorted, so sorting indices sorts keys) */
    size_t cap = 3 * (size_t)ncomp, nb = 0;
    long* be = (long*)malloc(sizeof(long) * (cap ? cap : 1));
    for (int t = 0; t < ncomp; ++t) {
        uint32_t f = comp[t];
        uint32_t v[3] = {faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]};
        for (int k = 0; k < 3; ++k) {
            uint32_t x = v[k], y = v[(k + 1) % 3];
            long e = edgetab_find(et, x < y ? x : y, x < y ? y : x);
            if (is_bnd[e]) be[nb++] = e;
        }
    }
    ch->nloops = 0; ch->loops = NULL; ch->loop_len = NULL;
    if (nb == 0) { free(be); return; }
    /* sort + dedupe edge indices */
    for (size_t i = 1; i < nb; ++i) { long x = be[i]; size_t j = i; while (j > 0 && be[j - 1] > x) { be[j] = be[j - 1]; j--; } be[j] = x; }
    size_t u = 0;
    for (size_t i = 0; i < nb; ++i) if (u == 0 || be[u - 1] != be[i]) be[u++] = be[i];
    nb = u;
    /* adjacency (BTreeMap<u32, Vec<u32>>): vertices sorted; per-vertex lists in edge-key order */
    uint32_t* verts = (uint32_t*)malloc(sizeof(uint32_t) * 2 * nb);
    for (size_t i = 0; i < nb; ++i) { verts[2 * i] = et->ka[be[i]]; verts[2 * i + 1] = et->kb[be[i]]; }
    qsort(verts, 2 * nb, sizeof(uint32_t), cmp_u32);
    size_t nvv = 0;
    for (size_t i = 0; i < 2 * nb; ++i) if (nvv == 0 || verts[nvv - 1] != verts[i]) verts[nvv++] = verts[i];
    vu32* adj = (vu32*)calloc(nvv, sizeof(vu32));
    for (size_t i = 0; i < nb; ++i) {
        uint32_t a = et->ka[be[i]], b = et->kb[be[i]];
        vu32_push(&adj[bsearch_u32(verts, nvv, a)], b);
        vu32_push(&adj[bsearch_u32(verts, nvv, b)], a);
    }
    /* used: parallel to be[] (sorted edge indices) */
    uint8_t* used = (uint8_t*)calloc(nb, 1);
    vu32 allv = {0, 0, 0};
    vu32 lens = {0, 0, 0};
    for (size_t s = 0; s < nvv; ++s) {
        uint32_t start = verts[s];
        vu32* ns = &adj[s];
        int has_unused = 0;
        for (size_t q = 0; q < ns->n; ++q) {
            uint32_t nb2 = ns->a[q];
            long e = edgetab_find(et, start < nb2 ? start : nb2, start < nb2 ? nb2 : start);
            /* find position of e in be */
            size_t lo = 0, hi = nb; while (lo < hi) { size_t m = (lo + hi) / 2; if (be[m] < e) lo = m + 1; else hi = m; }
            if (!used[lo]) { has_unused = 1; break; }
        }
        if (!has_unused) continue;
        vu32 lp = {0, 0, 0};
        vu32_push(&lp, start);
        uint32_t cur = start;
        /* first step: smallest candidate (cands sorted) */
        uint32_t first = ns->a[0];
        for (size_t q = 1; q < ns->n; ++q) if (ns->a[q] < first) first = ns->a[q];
        {
            long e = edgetab_find(et, cur < first ? cur : first, cur < first ? first : cur);
            size_t lo = 0, hi = nb; while (lo < hi) { size_t m = (lo + hi) / 2; if (be[m] < e) lo = m + 1; else hi = m; }
            used[lo] = 1;
        }
        uint32_t prev = cur;
        cur = first;
        vu32_push(&lp, cur);
        for (;;) {
            long ci = bsearch_u32(verts, nvv, cur);
            vu32* cn = &adj[ci];
            /* nxts: adjacency order, != prev, edge unused */
            uint32_t* nx = (uint32_t*)malloc(sizeof(uint32_t) * (cn->n ? cn->n : 1));
            size_t nn = 0;
            for (size_t q = 0; q < cn->n; ++q) {
                uint32_t c2 = cn->a[q];
                if (c2 == prev) continue;
                long e = edgetab_find(et, cur < c2 ? cur : c2, cur < c2 ? c2 : cur);
                size_t lo = 0, hi = nb; while (lo < hi) { size_t m = (lo + hi) / 2; if (be[m] < e) lo = m + 1; else hi = m; }
                if (used[lo]) continue;
                nx[nn++] = c2;
            }
            if (nn == 0) { free(nx); break; }
            v3 d_in = v3sub(P3(pos, cur), P3(pos, prev));
            double ln = v3len(d_in);
            uint32_t next;
            if (ln < 1e-30) {
                next = nx[0];
            } else {
                v3 din = v3div(d_in, ln);
                int have = 0; double bd = 0.0; uint32_t bn = nx[0];
                for (size_t q = 0; q < nn; ++q) {
                    double dot = v3dot(v3sub(P3(pos, nx[q]), P3(pos, cur)), din);
                    if (have && dot <= bd) continue;
                    have = 1; bd = dot; bn = nx[q];
                }
                next = bn;
            }
            free(nx);
            {
                long e = edgetab_find(et, cur < next ? cur : next, cur < next ? next : cur);
                size_t lo = 0, hi = nb; while (lo < hi) { size_t m = (lo + hi) / 2; if (be[m] < e) lo = m + 1; else hi = m; }
                used[lo] = 1;
            }
            prev = cur;
            cur = next;
            vu32_push(&lp, cur);
            if (cur == start) break;
        }
        if (lp.n >= 4) {
            for (size_t q = 0; q < lp.n; ++q) vu32_push(&allv, lp.a[q]);
            vu32_push(&lens, (uint32_t)lp.n);
        }
        free(lp.a);
    }
    ch->nloops = (int)lens.n;
    ch->loop_len = (int*)malloc(sizeof(int) * (lens.n ? lens.n : 1));
    ch->loops = (uint32_t**)malloc(sizeof(uint32_t*) * (lens.n ? lens.n : 1));
    size_t o = 0;
    for (size_t l = 0; l < lens.n; ++l) {
        ch->loop_len[l] = (int)lens.a[l];
        ch->loops[l] = (uint32_t*)malloc(sizeof(uint32_t) * lens.a[l]);
        memcpy(ch->loops[l], allv.a + o, sizeof(uint32_t) * lens.a[l]);
        o += lens.a[l];
    }
    free(allv.a); free(lens.a);
    for (size_t i = 0; i < nvv; ++i) free(adj[i].a);
    free(adj); free(used); free(verts); free(be);
}

/* segment.rs:130-223: cut edges with dihedral > angle, DFS face components, sorted vertex sets, border loops */
orc_charts* orc_segment(const double* pos, int nv, const uint32_t* faces, int nf, double angle_min_deg) {
    (void)nv;
    orc_charts* out = (orc_charts*)calloc(1, sizeof(orc_charts));
    edgetab et;
    edgetab_build(faces, nf, &et);
    /* to_radians: x * (PI / 180) */
    double th = angle_min_deg * (3.14159265358979323846 / 180.0);
    uint8_t* is_cut = (uint8_t*)calloc(et.ne ? et.ne : 1, 1);
    uint8_t* is_bnd = (uint8_t*)calloc(et.ne ? et.ne : 1, 1);
    int ncut = 0;
    for (size_t e = 0; e < et.ne; ++e) {
        size_t cnt = et.off[e + 1] - et.off[e];
        double ang;
        if (cnt == 1) ang = 0.0;
        else if (cnt == 2) {
            v3 n1 = face_normal_unit(pos, faces, et.faces[et.off[e]]);
            v3 n2 = face_normal_unit(pos, faces, et.faces[et.off[e] + 1]);
            double d = v3dot(n1, n2);
            if (d < -1.0) d = -1.0;
            if (d > 1.0) d = 1.0;
            ang = acos(d);
        } else ang = 3.14159265358979323846;
        if (ang > th) { is_cut[e] = 1; ncut++; }
        is_bnd[e] = (uint8_t)(is_cut[e] || cnt == 1);
    }
    out->ncut = ncut;
    out->cut = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (size_t)(ncut ? ncut : 1));
    { int k = 0; for (size_t e = 0; e < et.ne; ++e) if (is_cut[e]) { out->cut[2 * k] = et.ka[e]; out->cut[2 * k + 1] = et.kb[e]; k++; } }
    uint8_t* seen = (uint8_t*)calloc((size_t)(nf ? nf : 1), 1);
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nf ? nf : 1));
    uint32_t* comp = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nf ? nf : 1));
    int cap = 8;
    out->ch = (orc_chart*)calloc((size_t)cap, sizeof(orc_chart));
    for (int fi = 0; fi < nf; ++fi) {
        if (seen[fi]) continue;
        size_t sp = 0, nc = 0;
        stack[sp++] = (uint32_t)fi;
        seen[fi] = 1;
        while (sp > 0) {
            uint32_t f = stack[--sp];
            comp[nc++] = f;
            uint32_t v[3] = {faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]};
            for (int k = 0; k < 3; ++k) {
                uint32_t x = v[k], y = v[(k + 1) % 3];
                long e = edgetab_find(&et, x < y ? x : y, x < y ? y : x);
                if (is_cut[e]) continue;
                for (size_t q = et.off[e]; q < et.off[e + 1]; ++q) {
                    uint32_t g = et.faces[q];
                    if (!seen[g]) { seen[g] = 1; stack[sp++] = g; }
                }
            }
        }
        if (out->n == cap) { cap *= 2; out->ch = (orc_chart*)realloc(out->ch, sizeof(orc_chart) * (size_t)cap); }
        orc_chart* ch = &out->ch[out->n++];
        memset(ch, 0, sizeof(*ch));
        ch->nfaces = (int)nc;
        ch->face_ids = (uint32_t*)malloc(sizeof(uint32_t) * nc);
        memcpy(ch->face_ids, comp, sizeof(uint32_t) * nc);
        uint32_t* vs = (uint32_t*)malloc(sizeof(uint32_t) * 3 * nc);
        for (size_t t = 0; t < nc; ++t) for (int k = 0; k < 3; ++k) vs[3 * t + k] = faces[3 * comp[t] + k];
        qsort(vs, 3 * nc, sizeof(uint32_t), cmp_u32);
        size_t nvs = 0;
        for (size_t t = 0; t < 3 * nc; ++t) if (nvs == 0 || vs[nvs - 1] != vs[t]) vs[nvs++] = vs[t];
        ch->nverts = (int)nvs;
        ch->vertex_ids = vs;
        ch->tris = (uint32_t*)malloc(sizeof(uint32_t) * 3 * nc);
        for (size_t t = 0; t < nc; ++t)
            for (int k = 0; k < 3; ++k) ch->tris[3 * t + k] = (uint32_t)bsearch_u32(vs, nvs, faces[3 * comp[t] + k]);
        /* border vertices */
        uint32_t* bv = (uint32_t*)malloc(sizeof(uint32_t) * 6 * nc);
        size_t nbv = 0;
        for (size_t t = 0; t < nc; ++t) {
            uint32_t f = comp[t];
            uint32_t v[3] = {faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]};
            for (int k = 0; k < 3; ++k) {
                uint32_t x = v[k], y = v[(k + 1) % 3];
                long e = edgetab_find(&et, x < y ? x : y, x < y ? y : x);
                if (is_bnd[e]) { bv[nbv++] = x; bv[nbv++] = y; }
            }
        }
        qsort(bv, nbv, sizeof(uint32_t), cmp_u32);
        size_t nb2 = 0;
        for (size_t t = 0; t < nbv; ++t) if (nb2 == 0 || bv[nb2 - 1] != bv[t]) bv[nb2++] = bv[t];
        ch->nborder = (int)nb2;
        ch->border = bv;
        chart_border_loops(pos, faces, comp, (int)nc, &et, is_bnd, ch);
        double a = 0.0;
        for (size_t t = 0; t < nc; ++t) a += face_area3d(pos, faces, comp[t]);
        ch->area3d = a;
    }
    free(seen); free(stack); free(comp); free(is_cut); free(is_bnd);
    edgetab_free(&et);
    return out;
}

int orc_charts_count(const orc_charts* c) { return c->n; }
int orc_charts_ncut(const orc_charts* c) { return c->ncut; }
void orc_charts_cut(const orc_charts* c, uint32_t* out) { memcpy(out, c->cut, sizeof(uint32_t) * 2 * (size_t)c->ncut); }
/* sizes: [nfaces, nverts, nborder, nloops] */
void orc_chart_sizes(const orc_charts* c, int i, int* sizes) {
    const orc_chart* ch = &c->ch[i];
    sizes[0] = ch->nfaces; sizes[1] = ch->nverts; sizes[2] = ch->nborder; sizes[3] = ch->nloops;
}
void orc_chart_get(const orc_charts* c, int i, uint32_t* face_ids, uint32_t* vertex_ids, uint32_t* tris,
                   uint32_t* border, double* area3d) {
    const orc_chart* ch = &c->ch[i];
    memcpy(face_ids, ch->face_ids, sizeof(uint32_t) * (size_t)ch->nfaces);
    memcpy(vertex_ids, ch->vertex_ids, sizeof(uint32_t) * (size_t)ch->nverts);
    memcpy(tris, ch->tris, sizeof(uint32_t) * 3 * (size_t)ch->nfaces);
    memcpy(border, ch->border, sizeof(uint32_t) * (size_t)ch->nborder);
    *area3d = ch->area3d;
}
int orc_chart_loop_len(const orc_charts* c, int i, int l) { return c->ch[i].loop_len[l]; }
void orc_chart_loop(const orc_charts* c, int i, int l, uint32_t* out) {
    memcpy(out, c->ch[i].loops[l], sizeof(uint32_t) * (size_t)c->ch[i].loop_len[l]);
}
void orc_charts_free(orc_charts* c) {
    if (!c) return;
    for (int i = 0; i < c->n; ++i) {
        orc_chart* ch = &c->ch[i];
        free(ch->face_ids); free(ch->vertex_ids); free(ch->tris); free(ch->border);
        for (int l = 0; l < ch->nloops; ++l) free(ch->loops[l]);
        free(ch->loops); free(ch->loop_len);
    }
    free(c->ch); free(c->cut); free(c);
}

/* ------------------------------------------------------------------ eig3 / tangent frame */
/* lscm_eig.rs:13-66 (Jacobi rotations, largest off-diagonal first) */
void orc_eig3(const double m[9], double w[3], double vout[9]) {
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) a[i][j] = m[3 * i + j];
    for (int it = 0; it < 64; ++it) {
        int p = 0, q = 1;
        double max_off = fabs(a[0][1]);
        static const int pr[2][2] = {{0, 2}, {1, 2}};
        for (int t = 0; t < 2; ++t) {
            double x = fabs(a[pr[t][0]][pr[t][1]]);
            if (x > max_off) { max_off = x; p = pr[t][0]; q = pr[t][1]; }
        }
        if (max_off <= 1e-300) break;
        double apq = a[p][q], app = a[p][p], aqq = a[q][q];
        double theta = 0.5 * atan2(2.0 * apq, app - aqq);
        double c = cos(theta), s = sin(theta);
        for (int r = 0; r < 3; ++r) {
            if (r == p || r == q) continue;
            double ar_p = c * a[r][p] + s * a[r][q];
            double ar_q = -s * a[r][p] + c * a[r][q];
            a[r][p] = ar_p; a[p][r] = ar_p; a[r][q] = ar_q; a[q][r] = ar_q;
        }
        a[p][p] = c * c * app + 2.0 * s * c * apq + s * s * aqq;
        a[q][q] = s * s * app - 2.0 * s * c * apq + c * c * aqq;
        a[p][q] = 0.0; a[q][p] = 0.0;
        for (int k = 0; k < 3; ++k) {
            double vp = v[p][k], vq = v[q][k];
            v[p][k] = c * vp - s * vq;
            v[q][k] = s * vp + c * vq;
        }
    }
    w[0] = a[0][0]; w[1] = a[1][1]; w[2] = a[2][2];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) vout[3 * i + j] = v[i][j];
}

/* f64::total_cmp ordering key (sign-magnitude to two's complement) */
static int64_t total_key(double x) {
    int64_t b; memcpy(&b, &x, 8);
    b ^= (int64_t)(((uint64_t)(b >> 63)) >> 1);
    return b;
}

/* lscm.rs:152-183 PCA projection onto the two dominant covariance directions */
void orc_tangent_frame(const double* p3, int n, double* p2) {
    if (n == 0) return;
    v3 c = {0.0, 0.0, 0.0};
    for (int i = 0; i < n; ++i) c = v3add(c, P3(p3, (size_t)i));
    c = v3div(c, (double)n);
    double cm[9] = {0};
    for (int i = 0; i < n; ++i) {
        v3 d = v3sub(P3(p3, (size_t)i), c);
        double vi[3] = {d.x, d.y, d.z};
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) cm[3 * a + b] += vi[a] * vi[b];
    }
    double w[3], v[9];
    orc_eig3(cm, w, v);
    /* stable sort of [0,1,2] by w descending (total_cmp) */
    int idx[3] = {0, 1, 2};
    for (int i = 1; i < 3; ++i) {
        int x = idx[i], j = i;
        while (j > 0 && total_key(w[idx[j - 1]]) < total_key(w[x])) { idx[j] = idx[j - 1]; j--; }
        idx[j] = x;
    }
    v3 e1 = {v[0 * 3 + idx[0]], v[1 * 3 + idx[0]], v[2 * 3 + idx[0]]};
    v3 e2 = {v[0 * 3 + idx[1]], v[1 * 3 + idx[1]], v[2 * 3 + idx[1]]};
    for (int i = 0; i < n; ++i) {
        v3 d = v3sub(P3(p3, (size_t)i), c);
        p2[2 * i] = v3dot(d, e1);
        p2[2 * i + 1] = v3dot(d, e2);
    }
}

/* lscm.rs:187-204 */
void orc_rect_point(double f, double a, double b, double* out) {
    double p2 = 2.0 * (a + b);
    double t = fmod(f, 1.0) * p2;
    double x0 = a / 2.0, y0 = -b / 2.0;
    if (t < a) { out[0] = x0 - t; out[1] = y0; return; }
    t -= a;
    if (t < b) { out[0] = -a / 2.0; out[1] = y0 + t; return; }
    t -= b;
    if (t < a) { out[0] = -a / 2.0 + t; out[1] = b / 2.0; return; }
    t -= a;
    out[0] = a / 2.0; out[1] = b / 2.0 - t;
}

static int cmp_size_t(const void* a, const void* b) {
    size_t x = *(const size_t*)a, y = *(const size_t*)b; return x < y ? -1 : (x > y);
}
/* lscm.rs:210-247 greedy farthest-point anchors; writes up to n ids (sorted), returns count */
int orc_anchor_ids(const double* p3, int n3, int n, int64_t* out) {
    if (n3 <= n) { for (int i = 0; i < n3; ++i) out[i] = i; return n3; }
    double min_x = 1.7976931348623157e308;
    for (int i = 0; i < n3; ++i) { double x = p3[3 * i]; min_x = x < min_x ? x : min_x; }
    size_t* chosen = (size_t*)malloc(sizeof(size_t) * (size_t)n);
    int nc = 0;
    size_t st = 0;
    for (int i = 0; i < n3; ++i) if (p3[3 * i] == min_x) { st = (size_t)i; break; }
    chosen[nc++] = st;
    double* d = (double*)malloc(sizeof(double) * (size_t)n3);
    while (nc < n) {
        for (int i = 0; i < n3; ++i) {
            double m = 1.7976931348623157e308;
            for (int c = 0; c < nc; ++c) {
                double x = v3len(v3sub(P3(p3, (size_t)i), P3(p3, chosen[c])));
                m = x < m ? x : m; /* f64::min */
            }
            d[i] = m;
        }
        for (int c = 0; c < nc; ++c) d[chosen[c]] = -1.0;
        size_t bi = 0; double bd = -1.7976931348623157e308;
        for (int i = 0; i < n3; ++i) if (d[i] > bd) { bi = (size_t)i; bd = d[i]; }
        chosen[nc++] = bi;
    }
    qsort(chosen, (size_t)nc, sizeof(size_t), cmp_size_t);
    for (int i = 0; i < nc; ++i) out[i] = (int64_t)chosen[i];
    free(chosen); free(d);
    return nc;
}

/* ------------------------------------------------------------------ pins */
/* ordered map source vertex -> target (BTreeMap insert semantics: later insert overwrites) */
typedef struct { uint32_t* key; double* uv; size_t n, cap; } pinmap;
static void pinmap_insert(pinmap* m, uint32_t k, double u, double v) {
    size_t lo = 0, hi = m->n;
    while (lo < hi) { size_t mid = (lo + hi) / 2; if (m->key[mid] < k) lo = mid + 1; else hi = mid; }
    if (lo < m->n && m->key[lo] == k) { m->uv[2 * lo] = u; m->uv[2 * lo + 1] = v; return; }
    if (m->n == m->cap) {
        m->cap = m->cap ? 2 * m->cap : 16;
        m->key = (uint32_t*)realloc(m->key, sizeof(uint32_t) * m->cap);
        m->uv = (double*)realloc(m->uv, sizeof(double) * 2 * m->cap);
    }
    memmove(m->key + lo + 1, m->key + lo, sizeof(uint32_t) * (m->n - lo));
    memmove(m->uv + 2 * (lo + 1), m->uv + 2 * lo, sizeof(double) * 2 * (m->n - lo));
    m->key[lo] = k; m->uv[2 * lo] = u; m->uv[2 * lo + 1] = v; m->n++;
}

/* lscm.rs:280-413 border pass: rectangle targets by 3-D arc-length fraction */
static void border_targets(const double* pos, const orc_chart* ch, pinmap* pins) {
    int n = ch->nloops;
    double area3d = ch->area3d;
    double** fracs = (double**)malloc(sizeof(double*) * (size_t)n);
    double* perims = (double*)malloc(sizeof(double) * (size_t)n);
    for (int li = 0; li < n; ++li) {
        int L = ch->loop_len[li];
        const uint32_t* l = ch->loops[li];
        double* segs = (double*)malloc(sizeof(double) * (size_t)(L > 1 ? L - 1 : 1));
        for (int k = 0; k + 1 < L; ++k) segs[k] = v3len(v3sub(P3(pos, l[k + 1]), P3(pos, l[k])));
        double tot = -0.0;
        for (int k = 0; k + 1 < L; ++k) tot += segs[k];
        double total = tot > 1e-30 ? tot : 1e-30;
        double* fr = (double*)malloc(sizeof(double) * (size_t)(L > 0 ? L : 1));
        fr[0] = 0.0;
        for (int k = 0; k + 1 < L; ++k) fr[k + 1] = fr[k] + segs[k];
        for (int k = 0; k < L; ++k) fr[k] /= total;
        fracs[li] = fr;
        double ps = -0.0;
        for (int k = 0; k + 1 < L; ++k) ps += segs[k];
        perims[li] = ps;
        free(segs);
    }
    if (n == 2) {
        double mx = 0.0;
        for (int k = 0; k < 2; ++k) mx = (perims[k] > mx || mx != mx) ? perims[k] : mx; /* fold(0, f64::max) */
        if (fabs(perims[0] - perims[1]) <= 0.25 * mx) {
            double w = 0.5 * (perims[0] + perims[1]);
            double h = area3d / (w > 1e-30 ? w : 1e-30);
            int l0n = ch->loop_len[0], l1n = ch->loop_len[1];
            const uint32_t* l0 = ch->loops[0];
            const uint32_t* l1 = ch->loops[1];
            if (l0n > 1 && l0[0] == l0[l0n - 1]) l0n--;
            if (l1n > 1 && l1[0] == l1[l1n - 1]) l1n--;
            for (int k = 0; k < l0n; ++k) pinmap_insert(pins, l0[k], -w / 2.0 + w * fracs[0][k], h / 2.0);
            for (int k1 = 0; k1 < l1n; ++k1) {
                v3 p1 = P3(pos, l1[k1]);
                int best = 0; double best_d = INFINITY;
                for (int k0 = 0; k0 < l0n; ++k0) {
                    double d = v3len(v3sub(P3(pos, l0[k0]), p1));
                    if (d < best_d) { best_d = d; best = k0; }
                }
                pinmap_insert(pins, l1[k1], -w / 2.0 + w * fracs[0][best], -h / 2.0);
            }
            for (int li = 0; li < n; ++li) free(fracs[li]);
            free(fracs); free(perims);
            return;
        }
    }
    double* sa = (double*)malloc(sizeof(double) * (size_t)n);
    double* sb = (double*)malloc(sizeof(double) * (size_t)n);
    if (n == 1) {
        double p = perims[0], a = area3d;
        if (p * p >= 16.0 * a) {
            double s_ = p / 2.0;
            double q = s_ * s_ - 4.0 * a;
            double d = sqrt(q > 0.0 ? q : 0.0);
            sa[0] = (s_ + d) / 2.0; sb[0] = (s_ - d) / 2.0;
        } else {
            double s = sqrt(a);
            sa[0] = s; sb[0] = s;
        }
    } else {
        v3 c = {0.0, 0.0, 0.0};
        for (int t = 0; t < ch->nverts; ++t) c = v3add(c, P3(pos, ch->vertex_ids[t]));
        c = v3div(c, (double)(ch->nverts > 1 ? ch->nverts : 1));
        double* dists = (double*)malloc(sizeof(double) * (size_t)n);
        for (int li = 0; li < n; ++li) {
            double s = 0.0;
            for (int k = 0; k < ch->loop_len[li]; ++k) s += v3len(v3sub(P3(pos, ch->loops[li][k]), c));
            dists[li] = s / (double)(ch->loop_len[li] > 1 ? ch->loop_len[li] : 1);
        }
        int* order = (int*)malloc(sizeof(int) * (size_t)n);
        for (int i = 0; i < n; ++i) order[i] = i;
        for (int i = 1; i < n; ++i) { /* stable insertion sort by total_cmp */
            int x = order[i], j = i;
            while (j > 0 && total_key(dists[order[j - 1]]) > total_key(dists[x])) { order[j] = order[j - 1]; j--; }
            order[j] = x;
        }
        double denom = 0.0;
        for (int k = 0; k < n; ++k) {
            int li = order[k];
            double sign = ((n - 1 - k) % 2 == 0) ? 1.0 : -1.0;
            denom += sign * perims[li] * perims[li];
        }
        double ad = fabs(denom);
        double s = 4.0 * sqrt(area3d / (ad > 1e-30 ? ad : 1e-30));
        for (int k = 0; k < n; ++k) { int li = order[k]; double side = s * perims[li] / 4.0; sa[li] = side; sb[li] = side; }
        free(dists); free(order);
    }
    for (int li = 0; li < n; ++li) {
        double a = sa[li], b = sb[li];
        int L = ch->loop_len[li];
        const uint32_t* l = ch->loops[li];
        int closed = (L > 1 && l[0] == l[L - 1]);
        int nvl = closed ? L - 1 : L;
        /* tangent frame of the loop positions (closing repeat included) */
        double* lp3 = (double*)malloc(sizeof(double) * 3 * (size_t)L);
        double* lp2 = (double*)malloc(sizeof(double) * 2 * (size_t)L);
        for (int k = 0; k < L; ++k) memcpy(lp3 + 3 * k, pos + 3 * (size_t)l[k], sizeof(double) * 3);
        orc_tangent_frame(lp3, L, lp2);
        double sa3 = 0.0;
        for (int i = 0; i < L; ++i) {
            int j = (i + 1) % L;
            sa3 += lp2[2 * i] * lp2[2 * j + 1] - lp2[2 * i + 1] * lp2[2 * j];
        }
        sa3 *= 0.5;
        for (int k = 0; k < nvl; ++k) {
            double f = sa3 > 0.0 ? fmod(1.0 - fracs[li][k], 1.0) : fracs[li][k];
            double uv[2];
            orc_rect_point(f, a, b, uv);
            pinmap_insert(pins, l[k], uv[0], uv[1]);
        }
        free(lp3); free(lp2);
    }
    free(sa); free(sb);
    for (int li = 0; li < n; ++li) free(fracs[li]);
    free(fracs); free(perims);
}

/* ------------------------------------------------------------------ edge weights */
typedef struct { uint32_t i, j; double w; } ewrec;
typedef struct { uint32_t i, j; double c; uint32_t ord; } ewtmp;
static int cmp_ewtmp(const void* pa, const void* pb) {
    const ewtmp* a = (const ewtmp*)pa; const ewtmp* b = (const ewtmp*)pb;
    if (a->i != b->i) return a->i < b->i ? -1 : 1;
    if (a->j != b->j) return a->j < b->j ? -1 : 1;
    return a->ord < b->ord ? -1 : (a->ord > b->ord);
}
/* lscm.rs:252-274: per-edge sum of clamped opposite-angle cotangents, keyed (min,max), ascending keys.
 * Contributions to one key are summed in triangle order (BTreeMap entry +=). */
static size_t edge_weights(const double* p3, const uint32_t* tris, int ntri, ewrec** out) {
    size_t m = 3 * (size_t)ntri, k = 0;
    ewtmp* t = (ewtmp*)malloc(sizeof(ewtmp) * (m ? m : 1));
    for (int f = 0; f < ntri; ++f) {
        uint32_t i = tris[3 * f], j = tris[3 * f + 1], kk = tris[3 * f + 2];
        uint32_t trip[3][3] = {{i, j, kk}, {j, kk, i}, {kk, i, j}};
        for (int q = 0; q < 3; ++q) {
            uint32_t e0 = trip[q][0], e1 = trip[q][1], o = trip[q][2];
            v3 a1 = v3sub(P3(p3, e1), P3(p3, o));
            v3 a2 = v3sub(P3(p3, e0), P3(p3, o));
            double la = v3len(a1), lb = v3len(a2);
            if (la <= EPS_DEGENERATE || lb <= EPS_DEGENERATE) continue;
            double s = v3len(v3cross(a1, a2)) / (la * lb);
            if (s <= EPS_DEGENERATE) continue;
            double cot = (v3dot(a1, a2) / (la * lb)) / s;
            cot = cot > 0.0 ? cot : 0.0; /* f64::max(0.0); NaN -> 0 */
            t[k].i = e0 < e1 ? e0 : e1; t[k].j = e0 < e1 ? e1 : e0; t[k].c = cot; t[k].ord = (uint32_t)(3 * f + q); k++;
        }
    }
    qsort(t, k, sizeof(ewtmp), cmp_ewtmp);
    ewrec* r = (ewrec*)malloc(sizeof(ewrec) * (k ? k : 1));
    size_t ne = 0;
    for (size_t q = 0; q < k; ++q) {
        if (ne == 0 || r[ne - 1].i != t[q].i || r[ne - 1].j != t[q].j) {
            r[ne].i = t[q].i; r[ne].j = t[q].j; r[ne].w = 0.0; ne++;
        }
        r[ne - 1].w += t[q].c;
    }
    free(t);
    *out = r;
    return ne;
}

/* ------------------------------------------------------------------ Sparse (sparse.rs) */
typedef struct { int32_t* c; double* v; int32_t n, cap; } srow;
static void srow_reserve(srow* r, int32_t cap) {
    if (r->cap >= cap) return;
    int32_t nc = r->cap ? r->cap : 4;
    while (nc < cap) nc *= 2;
    r->c = (int32_t*)realloc(r->c, sizeof(int32_t) * (size_t)nc);
    r->v = (double*)realloc(r->v, sizeof(double) * (size_t)nc);
    r->cap = nc;
}
static int32_t srow_find(const srow* r, int32_t c) {
    int32_t lo = 0, hi = r->n;
    while (lo < hi) { int32_t m = (lo + hi) / 2; if (r->c[m] < c) lo = m + 1; else hi = m; }
    return lo;
}
/* Sparse::add (sparse.rs:47-52): skip exact zeros, sum duplicates into the (col -> val) map */
static void srow_add(srow* r, int32_t c, double v) {
    if (v == 0.0) return;
    int32_t p = srow_find(r, c);
    if (p < r->n && r->c[p] == c) { r->v[p] += v; return; }
    srow_reserve(r, r->n + 1);
    memmove(r->c + p + 1, r->c + p, sizeof(int32_t) * (size_t)(r->n - p));
    memmove(r->v + p + 1, r->v + p, sizeof(double) * (size_t)(r->n - p));
    r->c[p] = c; r->v[p] = 0.0 + v; r->n++;
}

typedef struct { int32_t n; srow* rows; } sparse;
static void sparse_init(sparse* s, int32_t n) { s->n = n; s->rows = (srow*)calloc((size_t)(n ? n : 1), sizeof(srow)); }
static void sparse_free(sparse* s) {
    for (int32_t i = 0; i < s->n; ++i) { free(s->rows[i].c); free(s->rows[i].v); }
    free(s->rows);
}

typedef struct { int32_t* a; int32_t n, cap; } vi32;
static void vi32_push(vi32* v, int32_t x) {
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 4; v->a = (int32_t*)realloc(v->a, sizeof(int32_t) * (size_t)v->cap); }
    v->a[v->n++] = x;
}

typedef struct { int32_t key_len; int32_t idx; } mkord;
static int cmp_mkord(const void* pa, const void* pb) {
    const mkord* a = (const mkord*)pa; const mkord* b = (const mkord*)pb;
    if (a->key_len != b->key_len) return a->key_len < b->key_len ? -1 : 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

/* sparse.rs:91-265: Markowitz-style (initial row length, index) symmetric order, partial pivoting,
 * exact-zero removal, unit-lower L / upper U, O(n) substitution over stored entries.
 * Returns 0 on success, -1 on singular/non-finite (None). */
static int sparse_solve(const sparse* s, const double* b, double* xo) {
    int32_t n = s->n;
    if (n == 0) return 0;
    mkord* ord = (mkord*)malloc(sizeof(mkord) * (size_t)n);
    for (int32_t c = 0; c < n; ++c) { ord[c].key_len = s->rows[c].n; ord[c].idx = c; }
    qsort(ord, (size_t)n, sizeof(mkord), cmp_mkord);
    int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* inv = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    for (int32_t i = 0; i < n; ++i) { perm[i] = ord[i].idx; inv[perm[i]] = i; }
    free(ord);
    /* permuted working rows, sorted by permuted column */
    srow* a = (srow*)calloc((size_t)n, sizeof(srow));
    vi32* colrows = (vi32*)calloc((size_t)n, sizeof(vi32)); /* column -> row positions that may hold it */
    for (int32_t i = 0; i < n; ++i) {
        const srow* src = &s->rows[perm[i]];
        srow_reserve(&a[i], src->n);
        for (int32_t q = 0; q < src->n; ++q) {
            int32_t cc = inv[src->c[q]];
            int32_t p = srow_find(&a[i], cc);
            memmove(a[i].c + p + 1, a[i].c + p, sizeof(int32_t) * (size_t)(a[i].n - p));
            memmove(a[i].v + p + 1, a[i].v + p, sizeof(double) * (size_t)(a[i].n - p));
            a[i].c[p] = cc; a[i].v[p] = src->v[q]; a[i].n++;
            vi32_push(&colrows[cc], i);
        }
    }
    /* factors: L rows (ascending k, by row position), U rows (diag first, then ascending) */
    srow* Lr = (srow*)calloc((size_t)n, sizeof(srow));
    srow* Ur = (srow*)calloc((size_t)n, sizeof(srow));
    int32_t* swk = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* swp = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t nsw = 0;
    uint8_t* hasdiag = (uint8_t*)calloc((size_t)n, 1);
    int32_t* mark = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    for (int32_t i = 0; i < n; ++i) mark[i] = -1;
    srow tmp = {0, 0, 0, 0};
    for (int32_t k = 0; k < n; ++k) {
        /* partial pivot (sparse.rs:139-153): scanning r = k.. ascending with a strict '>' keeps the smallest
         * row among equal maxima; the column list is unordered, so apply that rule explicitly */
        int32_t piv = k;
        double piv_val = 0.0;
        vi32* cl = &colrows[k];
        for (int32_t q = 0; q < cl->n; ++q) {
            int32_t r = cl->a[q];
            if (r < k) continue;
            int32_t p = srow_find(&a[r], k);
            if (!(p < a[r].n && a[r].c[p] == k)) continue;
            double v = fabs(a[r].v[p]);
            if (v > piv_val || (v == piv_val && v > 0.0 && r < piv)) { piv_val = v; piv = r; }
        }
        if (piv_val < 1e-300) continue;
        if (piv != k) {
            swk[nsw] = k; swp[nsw] = piv; nsw++;
            srow t = a[piv]; a[piv] = a[k]; a[k] = t;
            srow tl = Lr[piv]; Lr[piv] = Lr[k]; Lr[k] = tl;
            for (int32_t q = 0; q < a[k].n; ++q) vi32_push(&colrows[a[k].c[q]], k);
            for (int32_t q = 0; q < a[piv].n; ++q) vi32_push(&colrows[a[piv].c[q]], piv);
        }
        int32_t pk = srow_find(&a[k], k);
        double pivot = (pk < a[k].n && a[k].c[pk] == k) ? a[k].v[pk] : 0.0;
        hasdiag[k] = 1;
        srow_reserve(&Ur[k], a[k].n + 1);
        Ur[k].c[0] = k; Ur[k].v[0] = pivot; Ur[k].n = 1;
        int32_t ustart = srow_find(&a[k], k + 1);
        for (int32_t q = ustart; q < a[k].n; ++q) {
            if (a[k].v[q] != 0.0) { Ur[k].c[Ur[k].n] = a[k].c[q]; Ur[k].v[Ur[k].n] = a[k].v[q]; Ur[k].n++; }
        }
        /* eliminate rows r > k with a[r][k] != 0 */
        int32_t cnt = cl->n;
        for (int32_t q = 0; q < cnt; ++q) {
            int32_t r = cl->a[q];
            if (r <= k || mark[r] == k) continue;
            mark[r] = k;
            int32_t p = srow_find(&a[r], k);
            if (!(p < a[r].n && a[r].c[p] == k) || a[r].v[p] == 0.0) continue;
            double factor = a[r].v[p] / pivot;
            srow_reserve(&Lr[r], Lr[r].n + 1);
            Lr[r].c[Lr[r].n] = k; Lr[r].v[Lr[r].n] = factor; Lr[r].n++;
            /* merge: row r without col k, updated by pivot-row entries with col > k */
            srow_reserve(&tmp, a[r].n + (a[k].n - ustart));
            tmp.n = 0;
            int32_t x = 0, y = ustart;
            while (x < a[r].n || y < a[k].n) {
                int32_t cx = x < a[r].n ? a[r].c[x] : INT32_MAX;
                int32_t cy = y < a[k].n ? a[k].c[y] : INT32_MAX;
                if (cx == k) { x++; continue; }
                if (cx < cy) { tmp.c[tmp.n] = cx; tmp.v[tmp.n] = a[r].v[x]; tmp.n++; x++; }
                else {
                    double cur = 0.0;
                    if (cx == cy) { cur = a[r].v[x]; x++; }
                    double nv = cur - factor * a[k].v[y];
                    if (nv != 0.0) {
                        tmp.c[tmp.n] = cy; tmp.v[tmp.n] = nv; tmp.n++;
                        if (cx != cy) vi32_push(&colrows[cy], r);
                    }
                    y++;
                }
            }
            srow t = a[r]; a[r] = tmp; tmp = t;
        }
    }
    free(tmp.c); free(tmp.v);
    int rc = 0;
    /* solve (sparse.rs:219-265) */
    double* bp = (double*)malloc(sizeof(double) * (size_t)n);
    double* y = (double*)malloc(sizeof(double) * (size_t)n);
    double* x = (double*)malloc(sizeof(double) * (size_t)n);
    for (int32_t i = 0; i < n; ++i) bp[i] = b[perm[i]];
    for (int32_t q = 0; q < nsw; ++q) { double t = bp[swk[q]]; bp[swk[q]] = bp[swp[q]]; bp[swp[q]] = t; }
    for (int32_t i = 0; i < n; ++i) {
        double sum = bp[i];
        for (int32_t q = 0; q < Lr[i].n; ++q) sum -= Lr[i].v[q] * y[Lr[i].c[q]];
        y[i] = sum;
    }
    for (int32_t i = n - 1; i >= 0; --i) {
        double diag = hasdiag[i] ? Ur[i].v[0] : 0.0;
        if (fabs(diag) < 1e-300) { rc = -1; break; }
        double sum = y[i];
        for (int32_t q = 1; q < Ur[i].n; ++q) sum -= Ur[i].v[q] * x[Ur[i].c[q]];
        x[i] = sum / diag;
    }
    if (rc == 0) for (int32_t i = 0; i < n; ++i) if (!isfinite(x[i])) { rc = -1; break; }
    if (rc == 0) for (int32_t i = 0; i < n; ++i) xo[perm[i]] = x[i];
    for (int32_t i = 0; i < n; ++i) { free(a[i].c); free(a[i].v); free(Lr[i].c); free(Lr[i].v); free(Ur[i].c); free(Ur[i].v); free(colrows[i].a); }
    free(a); free(Lr); free(Ur); free(colrows); free(swk); free(swp); free(hasdiag); free(mark);
    free(bp); free(y); free(x); free(perm); free(inv);
    return rc;
}

/* Public: build a Sparse from COO (added in order) and solve. Returns 0 / -1. */
int orc_sparse_solve(int32_t n, int64_t ncoo, const int32_t* rows, const int32_t* cols, const double* vals,
                     const double* b, double* x) {
    sparse s;
    sparse_init(&s, n);
    for (int64_t q = 0; q < ncoo; ++q) srow_add(&s.rows[rows[q]], cols[q], vals[q]);
    int rc = sparse_solve(&s, b, x);
    sparse_free(&s);
    return rc;
}

/* ------------------------------------------------------------------ chart system (lscm.rs:512-666) */
typedef struct {
    int nv, nfree, npins;
    double* p3;       /* nv*3 chart-local positions */
    double* p2;       /* nv*2 tangent frame */
    int32_t* pin_ids; /* local ids, ascending source id */
    double* pin_uv;   /* npins*2 */
    int32_t* free_pos;/* nv, -1 for pins */
    int32_t* free_ids;/* nfree */
    sparse m;         /* (2*nfree+1)^2 system incl. slack */
    double* rhs;      /* 2*nfree+1 */
} chartsys;

static void chartsys_free(chartsys* cs) {
    free(cs->p3); free(cs->p2); free(cs->pin_ids); free(cs->pin_uv); free(cs->free_pos); free(cs->free_ids);
    free(cs->rhs);
    sparse_free(&cs->m);
}

/* pins (lscm.rs:524-542), edge weights (:545), free map (:582-592), assembly (:598-676) for the default
 * path (angle_mix = 1 -> B = 0, keep_metric = false) with a generic mix_b for completeness. */
static void chartsys_build(const double* pos, const orc_chart* ch, double angle_mix, double mix_w, chartsys* cs) {
    int nv = ch->nverts;
    cs->nv = nv;
    cs->p3 = (double*)malloc(sizeof(double) * 3 * (size_t)(nv ? nv : 1));
    for (int i = 0; i < nv; ++i) memcpy(cs->p3 + 3 * i, pos + 3 * (size_t)ch->vertex_ids[i], sizeof(double) * 3);
    cs->p2 = (double*)malloc(sizeof(double) * 2 * (size_t)(nv ? nv : 1));
    orc_tangent_frame(cs->p3, nv, cs->p2);
    pinmap pins = {0, 0, 0, 0};
    if (ch->nloops == 0) {
        int64_t an[4];
        int na = orc_anchor_ids(cs->p3, nv, 4, an);
        for (int q = 0; q < na; ++q) pinmap_insert(&pins, ch->vertex_ids[an[q]], cs->p2[2 * an[q]], cs->p2[2 * an[q] + 1]);
    } else {
        border_targets(pos, ch, &pins);
    }
    cs->pin_ids = (int32_t*)malloc(sizeof(int32_t) * (pins.n ? pins.n : 1));
    cs->pin_uv = (double*)malloc(sizeof(double) * 2 * (pins.n ? pins.n : 1));
    uint8_t* isfree = (uint8_t*)malloc((size_t)(nv ? nv : 1));
    memset(isfree, 1, (size_t)(nv ? nv : 1));
    double* pin_lookup = (double*)calloc(2 * (size_t)(nv ? nv : 1), sizeof(double));
    int np = 0;
    for (size_t q = 0; q < pins.n; ++q) {
        long li = bsearch_u32(ch->vertex_ids, (size_t)nv, pins.key[q]);
        if (li < 0) continue;
        int dup = 0;
        for (int t = 0; t < np; ++t) if (cs->pin_ids[t] == (int32_t)li) { dup = 1; break; }
        if (dup) continue;
        cs->pin_ids[np] = (int32_t)li;
        cs->pin_uv[2 * np] = pins.uv[2 * q]; cs->pin_uv[2 * np + 1] = pins.uv[2 * q + 1];
        isfree[li] = 0;
        pin_lookup[2 * li] = pins.uv[2 * q]; pin_lookup[2 * li + 1] = pins.uv[2 * q + 1];
        np++;
    }
    cs->npins = np;
    free(pins.key); free(pins.uv);
    ewrec* ew;
    size_t ne = edge_weights(cs->p3, ch->tris, ch->nfaces, &ew);
    cs->free_pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv ? nv : 1));
    cs->free_ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv ? nv : 1));
    int nfree = 0;
    for (int i = 0; i < nv; ++i) {
        if (isfree[i]) { cs->free_pos[i] = nfree; cs->free_ids[nfree++] = i; } else cs->free_pos[i] = -1;
    }
    cs->nfree = nfree;
    int32_t nvar = 2 * nfree + 1, aux = 2 * nfree;
    double mix_a = angle_mix, mix_b = 1.0 - mix_a;
    cs->rhs = (double*)calloc((size_t)nvar, sizeof(double));
    sparse_init(&cs->m, nvar);
    for (size_t e = 0; e < ne; ++e) {
        uint32_t i = ew[e].i, j = ew[e].j;
        double w = ew[e].w;
        if (w <= EPS_DEGENERATE) continue;
        v3 d3 = v3sub(P3(cs->p3, j), P3(cs->p3, i));
        double dd = v3dot(d3, d3);
        double c = w / (dd > EPS_DEGENERATE ? dd : EPS_DEGENERATE);
        double du = cs->p2[2 * j] - cs->p2[2 * i], dv = cs->p2[2 * j + 1] - cs->p2[2 * i + 1];
        for (int side = 0; side < 2; ++side) {
            uint32_t aa = side ? j : i, bb = side ? i : j;
            double su = side ? -du : du, sv = side ? -dv : dv;
            if (!isfree[aa]) continue;
            int32_t ra = 2 * cs->free_pos[aa];
            srow_add(&cs->m.rows[ra], ra, c);
            srow_add(&cs->m.rows[ra + 1], ra + 1, c);
            if (isfree[bb]) {
                int32_t col = 2 * cs->free_pos[bb];
                srow_add(&cs->m.rows[ra], col, -c);
                srow_add(&cs->m.rows[ra + 1], col + 1, -c);
                cs->rhs[ra] += mix_b * c * su;
                cs->rhs[ra + 1] += mix_b * c * sv;
            } else {
                cs->rhs[ra] += mix_b * c * su + c * pin_lookup[2 * bb];
                cs->rhs[ra + 1] += mix_b * c * sv + c * pin_lookup[2 * bb + 1];
            }
        }
    }
    srow_add(&cs->m.rows[aux], aux, mix_w);
    cs->rhs[aux] = mix_w;
    free(ew); free(isfree); free(pin_lookup);
}

/* Differentiable-form reference: the free block assembled from positions p3 with FROZEN pins (given ids and
 * targets, angle_mix = 1 so the rhs holds pin terms only), solved by the sparse LU (sparse.rs:219-265).
 * Writes uv (nv*2) and returns the loss E + sum_v |uv_v - target_v|^2 with E = sum over assembled edges of
 * c_e |uv_i - uv_j|^2 (c_e of lscm.rs:252-274,610).  *ok = 0 when the LU fails. */
double orc_lscm_loss_frozen(const double* p3, int nv, const uint32_t* tris, int ntri, const int32_t* pin_ids,
                            const double* pin_uv, int npins, const double* target, double* uv, int* ok) {
    uint8_t* isfree = (uint8_t*)malloc((size_t)(nv ? nv : 1));
    memset(isfree, 1, (size_t)(nv ? nv : 1));
    double* pin_lookup = (double*)calloc(2 * (size_t)(nv ? nv : 1), sizeof(double));
    for (int q = 0; q < npins; ++q) {
        isfree[pin_ids[q]] = 0;
        pin_lookup[2 * pin_ids[q]] = pin_uv[2 * q];
        pin_lookup[2 * pin_ids[q] + 1] = pin_uv[2 * q + 1];
    }
    int32_t* free_pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv ? nv : 1));
    int nfree = 0;
    for (int i = 0; i < nv; ++i) free_pos[i] = isfree[i] ? nfree++ : -1;
    ewrec* ew;
    size_t ne = edge_weights(p3, tris, ntri, &ew);
    int32_t nvar = 2 * nfree + 1, aux = 2 * nfree;
    double* rhs = (double*)calloc((size_t)nvar, sizeof(double));
    double* cval = (double*)calloc(ne ? ne : 1, sizeof(double));
    sparse m;
    sparse_init(&m, nvar);
    for (size_t e = 0; e < ne; ++e) {
        uint32_t i = ew[e].i, j = ew[e].j;
        if (ew[e].w <= EPS_DEGENERATE) continue;
        v3 d3 = v3sub(P3(p3, j), P3(p3, i));
        double dd = v3dot(d3, d3);
        double c = ew[e].w / (dd > EPS_DEGENERATE ? dd : EPS_DEGENERATE);
        cval[e] = c;
        for (int side = 0; side < 2; ++side) {
            uint32_t aa = side ? j : i, bb = side ? i : j;
            if (!isfree[aa]) continue;
            int32_t ra = 2 * free_pos[aa];
            srow_add(&m.rows[ra], ra, c);
            srow_add(&m.rows[ra + 1], ra + 1, c);
            if (isfree[bb]) {
                srow_add(&m.rows[ra], 2 * free_pos[bb], -c);
                srow_add(&m.rows[ra + 1], 2 * free_pos[bb] + 1, -c);
            } else {
                rhs[ra] += c * pin_lookup[2 * bb];
                rhs[ra + 1] += c * pin_lookup[2 * bb + 1];
            }
        }
    }
    srow_add(&m.rows[aux], aux, 1.0);
    rhs[aux] = 1.0;
    double* sol = (double*)malloc(sizeof(double) * (size_t)nvar);
    *ok = sparse_solve(&m, rhs, sol) == 0;
    double loss = 0.0;
    if (*ok) {
        for (int v = 0; v < nv; ++v) {
            uv[2 * v] = isfree[v] ? sol[2 * free_pos[v]] : pin_lookup[2 * v];
            uv[2 * v + 1] = isfree[v] ? sol[2 * free_pos[v] + 1] : pin_lookup[2 * v + 1];
        }
        for (size_t e = 0; e < ne; ++e) {
            if (cval[e] == 0.0) continue;
            double du = uv[2 * ew[e].i] - uv[2 * ew[e].j], dv = uv[2 * ew[e].i + 1] - uv[2 * ew[e].j + 1];
            loss += cval[e] * (du * du + dv * dv);
        }
        for (int v = 0; v < nv; ++v) {
            double du = uv[2 * v] - target[2 * v], dv = uv[2 * v + 1] - target[2 * v + 1];
            loss += du * du + dv * dv;
        }
    }
    sparse_free(&m);
    free(sol); free(rhs); free(cval); free(ew); free(free_pos); free(isfree); free(pin_lookup);
    return loss;
}

/* Gradient of the frozen-pin loss through the assembly, given the solution uv (nv*2, pins at their targets) and
 * the adjoint lam (nv*2, zero at pins; L lam = g_uv on the free vertices, L^T = L): for every assembled edge e =
 * (i, j) (w_e > EPS, c_e = w_e / max(|p_i - p_j|^2, EPS), lscm.rs:252-274,610)
 *   G_e = energy_weight |uv_i - uv_j|^2 - (lam_i - lam_j) . (uv_i - uv_j)
 * (dE/dc_e at fixed uv, and lam^T d(b - L x)/dc_e), and grad (nv*3) += G_e dc_e/dp: through |p_i - p_j|^2 and,
 * for every triangle corner o opposite e with cot > 0, through cot = (a1 . a2) / |a1 x a2| (a1 = p_e1 - p_o,
 * a2 = p_e0 - p_o; the corners that lscm.rs:252-274 skips or clamps to 0 contribute nothing). */
static int cmp_edge_key(const void* a, const void* b) {
    const ewrec* x = (const ewrec*)a; const ewrec* y = (const ewrec*)b;
    if (x->i != y->i) return x->i < y->i ? -1 : 1;
    if (x->j != y->j) return x->j < y->j ? -1 : 1;
    return 0;
}
void orc_lscm_grad_scatter(const double* p3, int nv, const uint32_t* tris, int ntri, const double* uv,
                           const double* lam, double energy_weight, double* grad) {
    ewrec* ew;
    size_t ne = edge_weights(p3, tris, ntri, &ew);
    double* G = (double*)calloc(ne ? ne : 1, sizeof(double));     /* dJ/dc_e */
    double* dd = (double*)calloc(ne ? ne : 1, sizeof(double));
    memset(grad, 0, sizeof(double) * 3 * (size_t)nv);
    for (size_t e = 0; e < ne; ++e) {
        uint32_t i = ew[e].i, j = ew[e].j;
        if (ew[e].w <= EPS_DEGENERATE) continue;
        v3 d3 = v3sub(P3(p3, j), P3(p3, i));
        double q = v3dot(d3, d3);
        dd[e] = q > EPS_DEGENERATE ? q : EPS_DEGENERATE;
        double du = uv[2 * i] - uv[2 * j], dv = uv[2 * i + 1] - uv[2 * j + 1];
        double lu = lam[2 * i] - lam[2 * j], lv = lam[2 * i + 1] - lam[2 * j + 1];
        G[e] = energy_weight * (du * du + dv * dv) - (lu * du + lv * dv);
        if (q > EPS_DEGENERATE) { /* d(w / |p_j - p_i|^2) through the length */
            double f = G[e] * 2.0 * ew[e].w / (q * q);
            grad[3 * j] -= f * d3.x; grad[3 * j + 1] -= f * d3.y; grad[3 * j + 2] -= f * d3.z;
            grad[3 * i] += f * d3.x; grad[3 * i + 1] += f * d3.y; grad[3 * i + 2] += f * d3.z;
        }
    }
    for (int t = 0; t < ntri; ++t) {
        uint32_t a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        uint32_t trip[3][3] = {{a, b, c}, {b, c, a}, {c, a, b}};
        for (int k = 0; k < 3; ++k) {
            uint32_t e0 = trip[k][0], e1 = trip[k][1], o = trip[k][2];
            v3 a1 = v3sub(P3(p3, e1), P3(p3, o));
            v3 a2 = v3sub(P3(p3, e0), P3(p3, o));
            double la = v3len(a1), lb = v3len(a2);
            if (la <= EPS_DEGENERATE || lb <= EPS_DEGENERATE) continue;
            v3 n = v3cross(a1, a2);
            double N = v3len(n);
            if (N / (la * lb) <= EPS_DEGENERATE) continue;
            double D = v3dot(a1, a2);
            if (!(D / N > 0.0)) continue; /* clamped to 0 */
            ewrec key;
            key.i = e0 < e1 ? e0 : e1; key.j = e0 < e1 ? e1 : e0;
            ewrec* hit = (ewrec*)bsearch(&key, ew, ne, sizeof(ewrec), cmp_edge_key);
            if (!hit || hit->w <= EPS_DEGENERATE) continue;
            size_t e = (size_t)(hit - ew);
            double gw = G[e] / dd[e]; /* dJ/dw_e */
            v3 nh = {n.x / N, n.y / N, n.z / N};
            v3 c1 = v3cross(a2, nh), c2 = v3cross(nh, a1); /* dN/da1, dN/da2 */
            double N2 = N * N;
            v3 g1 = {(a2.x * N - D * c1.x) / N2, (a2.y * N - D * c1.y) / N2, (a2.z * N - D * c1.z) / N2};
            v3 g2 = {(a1.x * N - D * c2.x) / N2, (a1.y * N - D * c2.y) / N2, (a1.z * N - D * c2.z) / N2};
            grad[3 * e1] += gw * g1.x; grad[3 * e1 + 1] += gw * g1.y; grad[3 * e1 + 2] += gw * g1.z;
            grad[3 * e0] += gw * g2.x; grad[3 * e0 + 1] += gw * g2.y; grad[3 * e0 + 2] += gw * g2.z;
            grad[3 * o] -= gw * (g1.x + g2.x); grad[3 * o + 1] -= gw * (g1.y + g2.y);
            grad[3 * o + 2] -= gw * (g1.z + g2.z);
        }
    }
    free(G); free(dd); free(ew);
}

/* The frozen-pin VJP by the sparse LU (sparse.rs:219-265): uv from the forward solve, lam from a second solve of
 * the same matrix with the free rows of g_uv as the right-hand side (the slack row 0), then
 * orc_lscm_grad_scatter.  uv_out (nv*2), grad (nv*3); returns 0 on success, -1 when an LU solve fails. */
int orc_lscm_vjp_frozen(const double* p3, int nv, const uint32_t* tris, int ntri, const int32_t* pin_ids,
                        const double* pin_uv, int npins, const double* g_uv, double energy_weight, double* uv_out,
                        double* grad) {
    double* target = (double*)calloc(2 * (size_t)(nv ? nv : 1), sizeof(double));
    int ok = 0;
    (void)orc_lscm_loss_frozen(p3, nv, tris, ntri, pin_ids, pin_uv, npins, target, uv_out, &ok);
    free(target);
    if (!ok) return -1;
    uint8_t* isfree = (uint8_t*)malloc((size_t)(nv ? nv : 1));
    memset(isfree, 1, (size_t)(nv ? nv : 1));
    for (int q = 0; q < npins; ++q) isfree[pin_ids[q]] = 0;
    int32_t* free_pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv ? nv : 1));
    int nfree = 0;
    for (int i = 0; i < nv; ++i) free_pos[i] = isfree[i] ? nfree++ : -1;
    ewrec* ew;
    size_t ne = edge_weights(p3, tris, ntri, &ew);
    int32_t nvar = 2 * nfree + 1, aux = 2 * nfree;
    sparse m;
    sparse_init(&m, nvar);
    for (size_t e = 0; e < ne; ++e) { /* the matrix of orc_lscm_loss_frozen */
        uint32_t i = ew[e].i, j = ew[e].j;
        if (ew[e].w <= EPS_DEGENERATE) continue;
        v3 d3 = v3sub(P3(p3, j), P3(p3, i));
        double dd = v3dot(d3, d3);
        double c = ew[e].w / (dd > EPS_DEGENERATE ? dd : EPS_DEGENERATE);
        for (int side = 0; side < 2; ++side) {
            uint32_t aa = side ? j : i, bb = side ? i : j;
            if (!isfree[aa]) continue;
            int32_t ra = 2 * free_pos[aa];
            srow_add(&m.rows[ra], ra, c);
            srow_add(&m.rows[ra + 1], ra + 1, c);
            if (isfree[bb]) {
                srow_add(&m.rows[ra], 2 * free_pos[bb], -c);
                srow_add(&m.rows[ra + 1], 2 * free_pos[bb] + 1, -c);
            }
        }
    }
    srow_add(&m.rows[aux], aux, 1.0);
    double* rhs = (double*)calloc((size_t)nvar, sizeof(double));
    double* sol = (double*)malloc(sizeof(double) * (size_t)nvar);
    for (int v = 0; v < nv; ++v)
        if (isfree[v]) { rhs[2 * free_pos[v]] = g_uv[2 * v]; rhs[2 * free_pos[v] + 1] = g_uv[2 * v + 1]; }
    int rc = sparse_solve(&m, rhs, sol) == 0 ? 0 : -1;
    if (rc == 0) {
        double* lam = (double*)calloc(2 * (size_t)(nv ? nv : 1), sizeof(double));
        for (int v = 0; v < nv; ++v)
            if (isfree[v]) { lam[2 * v] = sol[2 * free_pos[v]]; lam[2 * v + 1] = sol[2 * free_pos[v] + 1]; }
        orc_lscm_grad_scatter(p3, nv, tris, ntri, uv_out, lam, energy_weight, grad);
        free(lam);
    }
    sparse_free(&m);
    free(rhs); free(sol); free(ew); free(free_pos); free(isfree);
    return rc;
}

/* Export the chart system: sizes[0]=nv, [1]=nfree, [2]=npins, [3]=nnz of the leading 2*nfree block.
 * Call with NULL arrays first to get sizes. CSR follows Sparse::to_csr(2*nfree) (sparse.rs:69-84). */
int orc_chart_system(const orc_charts* cc, int ci, const double* pos, double angle_mix, double mix_w,
                     int64_t* sizes, double* p2, int32_t* pin_ids, double* pin_uv, int32_t* free_ids,
                     double* vals, int32_t* cols, int32_t* row_ptr, double* rhs) {
    chartsys cs;
    chartsys_build(pos, &cc->ch[ci], angle_mix, mix_w, &cs);
    int32_t lim = 2 * cs.nfree;
    int64_t nnz = 0;
    for (int32_t r = 0; r < lim; ++r)
        for (int32_t q = 0; q < cs.m.rows[r].n; ++q) if (cs.m.rows[r].c[q] < lim) nnz++;
    sizes[0] = cs.nv; sizes[1] = cs.nfree; sizes[2] = cs.npins; sizes[3] = nnz;
    if (p2) memcpy(p2, cs.p2, sizeof(double) * 2 * (size_t)cs.nv);
    if (pin_ids) memcpy(pin_ids, cs.pin_ids, sizeof(int32_t) * (size_t)cs.npins);
    if (pin_uv) memcpy(pin_uv, cs.pin_uv, sizeof(double) * 2 * (size_t)cs.npins);
    if (free_ids) memcpy(free_ids, cs.free_ids, sizeof(int32_t) * (size_t)cs.nfree);
    if (vals && cols && row_ptr) {
        int64_t k = 0;
        row_ptr[0] = 0;
        for (int32_t r = 0; r < lim; ++r) {
            for (int32_t q = 0; q < cs.m.rows[r].n; ++q)
                if (cs.m.rows[r].c[q] < lim) { vals[k] = cs.m.rows[r].v[q]; cols[k] = cs.m.rows[r].c[q]; k++; }
            row_ptr[r + 1] = (int32_t)k;
        }
    }
    if (rhs) memcpy(rhs, cs.rhs, sizeof(double) * (size_t)(lim + 1));
    chartsys_free(&cs);
    return 0;
}

/* ------------------------------------------------------------------ unfold driver (lscm.rs:512-727) */
/* CPU backend only (the oracle path). solver: 0 = sparse LU (SolverBackend::Cpu). uv_out: nv*2.
 * stats[0] = iters, stats[1] = err_accum, stats[2] = target, stats[3] = area3d. Returns nv. */
int orc_unfold_chart(const orc_charts* cc, int ci, const double* pos, int max_iter_opt, double angle_mix,
                     double mix_w, double* uv_out, double* stats) {
    const orc_chart* ch = &cc->ch[ci];
    chartsys cs;
    chartsys_build(pos, ch, angle_mix, mix_w, &cs);
    int nv = cs.nv;
    double* uv = (double*)malloc(sizeof(double) * 2 * (size_t)(nv ? nv : 1));
    memcpy(uv, cs.p2, sizeof(double) * 2 * (size_t)nv);
    double err_accum = 1.0, step = 1.0, prev_err = INFINITY;
    int iters_done = 0;
    int max_iter = max_iter_opt > 0 ? max_iter_opt : DEFAULT_MAX_ITER;
    int32_t nvar = 2 * cs.nfree + 1;
    double* sol = (double*)malloc(sizeof(double) * (size_t)nvar);
    double* new_uv = (double*)malloc(sizeof(double) * 2 * (size_t)(nv ? nv : 1));
    double* dvv = (double*)malloc(sizeof(double) * 2 * (size_t)(nv ? nv : 1));
    for (int it = 0; it < max_iter; ++it) {
        iters_done = it + 1;
        /* the system is re-assembled every iteration; with B = 0 and no area row it is identical, and the
         * LU of a fresh Sparse gives the identical solution, so one assembly serves every iteration */
        if (sparse_solve(&cs.m, cs.rhs, sol) != 0) break;
        memcpy(new_uv, uv, sizeof(double) * 2 * (size_t)nv);
        for (int q = 0; q < cs.nfree; ++q) {
            int i = cs.free_ids[q];
            new_uv[2 * i] = sol[2 * q]; new_uv[2 * i + 1] = sol[2 * q + 1];
        }
        for (int q = 0; q < cs.npins; ++q) {
            new_uv[2 * cs.pin_ids[q]] = cs.pin_uv[2 * q]; new_uv[2 * cs.pin_ids[q] + 1] = cs.pin_uv[2 * q + 1];
        }
        double sdv = -0.0, suv = -0.0;
        for (int i = 0; i < nv; ++i) {
            dvv[2 * i] = new_uv[2 * i] - uv[2 * i]; dvv[2 * i + 1] = new_uv[2 * i + 1] - uv[2 * i + 1];
            sdv += dvv[2 * i] * dvv[2 * i] + dvv[2 * i + 1] * dvv[2 * i + 1];
        }
        for (int i = 0; i < nv; ++i) suv += uv[2 * i] * uv[2 * i] + uv[2 * i + 1] * uv[2 * i + 1];
        double norm_dv = sqrt(sdv), norm_uv = sqrt(suv);
        double err = norm_dv / (norm_uv > 1e-30 ? norm_uv : 1e-30);
        double applied = err * step;
        if (prev_err > 1e-30 && err > 2.0 * prev_err) { double s2 = step * 0.5; step = s2 > 0.05 ? s2 : 0.05; }
        else if (it > 3) { double s2 = step * 1.2; step = s2 < 1.0 ? s2 : 1.0; }
        if (applied < 1e-8 || (prev_err > 1e-30 && it != 0 && applied / prev_err <= STAGNATION)) break;
        for (int i = 0; i < nv; ++i) { uv[2 * i] = uv[2 * i] + dvv[2 * i] * step; uv[2 * i + 1] = uv[2 * i + 1] + dvv[2 * i + 1] * step; }
        prev_err = applied;
        err_accum += err;
        if (max_iter * DEFAULT_MAX_ITER_GUARD < it) break;
    }
    memcpy(uv_out, uv, sizeof(double) * 2 * (size_t)nv);
    stats[0] = iters_done; stats[1] = err_accum; stats[2] = ch->area3d; stats[3] = ch->area3d;
    free(uv); free(sol); free(new_uv); free(dvv);
    chartsys_free(&cs);
    return nv;
}

/* Worker-thread path (pipeline.rs:171-194 + exec.rs:28-60): contiguous chunks, results in input order.
 * Unfolds charts [0, nch) with `workers` threads; uv written at uv_off[ci]*2. */
typedef struct {
    const orc_charts* cc; const double* pos; int lo, hi; double angle_mix, mix_w; double* uv; const int64_t* off;
    int max_iter;
} parjob;
static void* par_worker(void* arg) {
    parjob* j = (parjob*)arg;
    double st[4];
    for (int ci = j->lo; ci < j->hi; ++ci)
        orc_unfold_chart(j->cc, ci, j->pos, j->max_iter, j->angle_mix, j->mix_w, j->uv + 2 * j->off[ci], st);
    return NULL;
}
/* max_iter 0: the driver's own iteration count (two factor+solve passes with the default options); 1: one
 * factor+solve per chart, the unit of "unknowns solved". */
int orc_par_unfold_iters(const orc_charts* cc, const double* pos, int workers, int max_iter, double angle_mix,
                         double mix_w, const int64_t* uv_off, double* uv) {
    int n = cc->n;
    if (workers <= 1 || n <= 1) {
        parjob j = {cc, pos, 0, n, angle_mix, mix_w, uv, uv_off, max_iter};
        par_worker(&j);
        return 0;
    }
    if (workers > n) workers = n;
    int chunk = (n + workers - 1) / workers;
    pthread_t th[256];
    parjob jobs[256];
    int nt = 0;
    for (int lo = 0; lo < n && nt < 256; lo += chunk) {
        int hi = lo + chunk < n ? lo + chunk : n;
        parjob j = {cc, pos, lo, hi, angle_mix, mix_w, uv, uv_off, max_iter};
        jobs[nt] = j;
        pthread_create(&th[nt], NULL, par_worker, &jobs[nt]);
        nt++;
    }
    for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    return 0;
}
int orc_par_unfold(const orc_charts* cc, const double* pos, int workers, double angle_mix, double mix_w,
                   const int64_t* uv_off, double* uv) {
    return orc_par_unfold_iters(cc, pos, workers, 0, angle_mix, mix_w, uv_off, uv);
}

/* ------------------------------------------------------------------ CG (lscm.rs:417-459) */
int orc_cg_solve(int32_t n, const double* vals, const int32_t* cols, const int32_t* row_ptr, const double* rhs,
                 int64_t max_iter, double* x) {
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    double* p = (double*)malloc(sizeof(double) * (size_t)n);
    double* ap = (double*)malloc(sizeof(double) * (size_t)n);
    double rz = -0.0, bb = -0.0;
    for (int32_t i = 0; i < n; ++i) { x[i] = 0.0; r[i] = rhs[i]; p[i] = rhs[i]; }
    for (int32_t i = 0; i < n; ++i) rz += r[i] * r[i];
    for (int32_t i = 0; i < n; ++i) bb += rhs[i] * rhs[i];
    double tol = sqrt(bb) * 1e-12 + 1e-30, tol2 = tol * tol;
    int rc = 0;
    for (int64_t it = 0; it < max_iter; ++it) {
        if (rz < tol2) break;
        for (int32_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) s += vals[k] * p[cols[k]];
            ap[i] = s;
        }
        double pap = -0.0;
        for (int32_t i = 0; i < n; ++i) pap += p[i] * ap[i];
        if (fabs(pap) < 1e-300) { rc = -1; break; }
        double alpha = rz / pap;
        for (int32_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * ap[i]; }
        double rzn = -0.0;
        for (int32_t i = 0; i < n; ++i) rzn += r[i] * r[i];
        if (rzn < tol2) break;
        for (int32_t i = 0; i < n; ++i) p[i] = r[i] + (rzn / rz) * p[i];
        rz = rzn;
    }
    if (rc == 0) for (int32_t i = 0; i < n; ++i) if (!isfinite(x[i])) { rc = -1; break; }
    free(r); free(p); free(ap);
    return rc;
}

/* ------------------------------------------------------------------ IC(0) (precond.rs:145-290) */
typedef struct { int32_t n; int64_t nnz; double* v; int32_t* c; int32_t* p; } csr;

/* factor_ic0: L with A's strict-lower pattern (ascending), diagonal LAST, pivot floor 1e-12*|a_ii|.
 * Outputs l_* (nnz_l = count of strict-lower + n); returns nnz_l or -1 (None). Call with NULL outputs
 * for the size. */
int64_t orc_ic0_factor(int32_t n, const double* vals, const int32_t* cols, const int32_t* row_ptr,
                       double* l_vals, int32_t* l_cols, int32_t* l_ptr) {
    int64_t nnzl = n;
    for (int32_t i = 0; i < n; ++i) for (int32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) if (cols[k] < i) nnzl++;
    if (!l_vals) return nnzl;
    /* rows[i]: strict lower entries of L sorted by column (dense-addressable through a scatter map) */
    int32_t* rp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t* rc_ = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnzl ? nnzl : 1));
    double* rv = (double*)malloc(sizeof(double) * (size_t)(nnzl ? nnzl : 1));
    double* diag = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
    int64_t w = 0;
    int64_t kk = 0;
    l_ptr[0] = 0;
    int32_t* lc = (int32_t*)malloc(sizeof(int32_t) * 64);
    double* la = (double*)malloc(sizeof(double) * 64);
    int lcap = 64;
    for (int32_t i = 0; i < n; ++i) {
        rp[i] = (int32_t)w;
        int cnt = 0;
        for (int32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) if (cols[k] < i) cnt++;
        if (cnt > lcap) { lcap = cnt; lc = (int32_t*)realloc(lc, sizeof(int32_t) * (size_t)lcap); la = (double*)realloc(la, sizeof(double) * (size_t)lcap); }
        int m = 0;
        for (int32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) if (cols[k] < i) { lc[m] = cols[k]; la[m] = vals[k]; m++; }
        for (int a = 1; a < m; ++a) { /* sort by column (stable) */
            int32_t c0 = lc[a]; double v0 = la[a]; int b = a;
            while (b > 0 && lc[b - 1] > c0) { lc[b] = lc[b - 1]; la[b] = la[b - 1]; b--; }
            lc[b] = c0; la[b] = v0;
        }
        for (int t = 0; t < m; ++t) {
            int32_t k = lc[t];
            double s = la[t];
            /* sum over j < k present in row i (so far) and row k, ascending j */
            int32_t qi = rp[i], qe = (int32_t)w, qk = rp[k], qke = rp[k + 1];
            while (qi < qe && rc_[qi] < k) {
                int32_t j = rc_[qi];
                while (qk < qke && rc_[qk] < j) qk++;
                if (qk < qke && rc_[qk] == j) s -= rv[qi] * rv[qk];
                qi++;
            }
            double l_kk = diag[k];
            if (fabs(l_kk) < 1e-300) { free(rp); free(rc_); free(rv); free(diag); free(lc); free(la); return -1; }
            rc_[w] = k; rv[w] = s / l_kk; w++;
        }
        rp[i + 1] = (int32_t)w;
        double a_ii = 0.0;
        for (int32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) if (cols[k] == i) { a_ii = vals[k]; break; }
        if (fabs(a_ii) < 1e-300) { free(rp); free(rc_); free(rv); free(diag); free(lc); free(la); return -1; }
        double sum = a_ii;
        for (int32_t q = rp[i]; q < rp[i + 1]; ++q) sum -= rv[q] * rv[q];
        double fa = fabs(a_ii);
        double floor_ = 1e-12 * (fa > 1e-30 ? fa : 1e-30);
        if (sum < floor_) sum = floor_;
        diag[i] = sqrt(sum);
        for (int32_t q = rp[i]; q < rp[i + 1]; ++q) { l_vals[kk] = rv[q]; l_cols[kk] = rc_[q]; kk++; }
        l_vals[kk] = diag[i]; l_cols[kk] = i; kk++;
        l_ptr[i + 1] = (int32_t)kk;
    }
    free(rp); free(rc_); free(rv); free(diag); free(lc); free(la);
    return nnzl;
}

/* tri_levels (precond.rs:227-249): level of each row; ascending order for L, descending for L^T. */
int32_t orc_tri_levels(int32_t n, const int32_t* cols, const int32_t* ptr, int descending, int32_t* level) {
    int32_t nl = 0;
    for (int32_t t = 0; t < n; ++t) {
        int32_t i = descending ? n - 1 - t : t;
        int32_t lv = 0;
        for (int32_t k = ptr[i]; k < ptr[i + 1] - 1; ++k) { int32_t d = level[cols[k]] + 1; lv = lv > d ? lv : d; }
        level[i] = lv;
        if (lv + 1 > nl) nl = lv + 1;
    }
    return nl;
}

/* IC(0) apply on the host (precond.rs:907-922): forward over L then backward over L^T (given as CSR, diag last) */
void orc_ic0_apply(int32_t n, const double* l_vals, const int32_t* l_cols, const int32_t* l_ptr, const double* lt_vals,
                   const int32_t* lt_cols, const int32_t* lt_ptr, const double* r, double* z) {
    double* y = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
    for (int32_t i = 0; i < n; ++i) {
        double s = r[i];
        for (int32_t k = l_ptr[i]; k < l_ptr[i + 1] - 1; ++k) s -= l_vals[k] * y[l_cols[k]];
        y[i] = s / l_vals[l_ptr[i + 1] - 1];
    }
    for (int32_t i = n - 1; i >= 0; --i) {
        double s = y[i];
        for (int32_t k = lt_ptr[i]; k < lt_ptr[i + 1] - 1; ++k) s -= lt_vals[k] * z[lt_cols[k]];
        z[i] = s / lt_vals[lt_ptr[i + 1] - 1];
    }
    free(y);
}

/* transpose_lower (precond.rs:253-290): rows ascending by column, diagonal last */
void orc_transpose_lower(int32_t n, const double* vals, const int32_t* cols, const int32_t* ptr, double* ov,
                         int32_t* oc, int32_t* op) {
    int32_t* cnt = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    for (int32_t i = 0; i < n; ++i) for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) cnt[cols[k] + 1]++;
    for (int32_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    memcpy(cur, cnt, sizeof(int32_t) * ((size_t)n + 1));
    /* entries of transposed row j: (i, v) for L(i, j); i ascending because rows are scanned ascending, and the
     * diagonal (i == j) is the smallest i, so it is moved last */
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
            int32_t j = cols[k];
            oc[cur[j]] = i; ov[cur[j]] = vals[k]; cur[j]++;
        }
    for (int32_t j = 0; j < n; ++j) {
        int32_t a = cnt[j], b = cnt[j + 1];
        /* diagonal is at position a (smallest row index in column j is j itself) */
        int32_t dc = oc[a]; double dv = ov[a];
        memmove(oc + a, oc + a + 1, sizeof(int32_t) * (size_t)(b - a - 1));
        memmove(ov + a, ov + a + 1, sizeof(double) * (size_t)(b - a - 1));
        oc[b - 1] = dc; ov[b - 1] = dv;
        op[j] = a;
    }
    op[n] = cnt[n];
    free(cnt); free(cur);
}

/* graph_coloring (precond.rs:52-77): greedy distance-1, index order, smallest free colour */
void orc_coloring(int32_t n, const int32_t* cols, const int32_t* row_ptr, uint32_t* color) {
    uint32_t* used = (uint32_t*)malloc(sizeof(uint32_t) * 64);
    int cap = 64;
    for (int32_t v = 0; v < n; ++v) color[v] = UINT32_MAX;
    for (int32_t v = 0; v < n; ++v) {
        int m = 0;
        for (int32_t k = row_ptr[v]; k < row_ptr[v + 1]; ++k) {
            int32_t c = cols[k];
            if (c != v && color[c] != UINT32_MAX) {
                if (m == cap) { cap *= 2; used = (uint32_t*)realloc(used, sizeof(uint32_t) * (size_t)cap); }
                used[m++] = color[c];
            }
        }
        qsort(used, (size_t)m, sizeof(uint32_t), cmp_u32);
        uint32_t k = 0;
        int prevset = 0; uint32_t prev = 0;
        for (int t = 0; t < m; ++t) {
            if (prevset && used[t] == prev) continue;
            prevset = 1; prev = used[t];
            if (used[t] == k) k++; else break;
        }
        color[v] = k;
    }
    free(used);
}

/* ------------------------------------------------------------------ post-solve chart metrics */
/* tri_signed (metrics.rs:31-34): signed doubled area of one UV triangle */
static double tri_signed2(v2 u, v2 v, v2 w) { return (v.u - u.u) * (w.v - u.v) - (w.u - u.u) * (v.v - u.v); }
static v2 UV(const double* uv, size_t i) { v2 r = {uv[2 * i], uv[2 * i + 1]}; return r; }

/* winding_flips (metrics.rs:39-56): triangles whose UV winding disagrees with the signed-area majority */
int64_t orc_winding_flips(const double* uv, const uint32_t* tris, int64_t ntri) {
    if (ntri == 0) return 0;
    double sum = 0.0;
    for (int64_t t = 0; t < ntri; ++t)
        sum += tri_signed2(UV(uv, tris[3 * t]), UV(uv, tris[3 * t + 1]), UV(uv, tris[3 * t + 2]));
    const double dom = (sum < 0.0) ? -1.0 : 1.0;
    int64_t n = 0;
    for (int64_t t = 0; t < ntri; ++t)
        if (tri_signed2(UV(uv, tris[3 * t]), UV(uv, tris[3 * t + 1]), UV(uv, tris[3 * t + 2])) * dom <= 0.0) ++n;
    return n;
}

/* svd2 (metrics.rs:59-71): singular values of a 2x2 matrix through the eigenvalues of M^T M */
static void svd2(const double m[2][2], double* s1, double* s2) {
    const double a = m[0][0] * m[0][0] + m[1][0] * m[1][0];
    const double b = m[0][0] * m[0][1] + m[1][0] * m[1][1];
    const double c = m[0][1] * m[0][1] + m[1][1] * m[1][1];
    const double tr = a + c;
    const double d = 0.5 * (a - c);
    const double disc = sqrt(fmax(d * d + b * b, 0.0));
    const double l1 = 0.5 * tr + disc;
    const double l2 = fmax(0.5 * tr - disc, 0.0);
    *s1 = sqrt(fmax(l1, 0.0));
    *s2 = sqrt(l2);
}

/* chart_distortion (metrics.rs:73-141).  p3: the chart's vertex positions in chart-local order (the mesh
 * positions of vertex_ids), tris: chart-local triangles (the mesh faces' corners in face order), uv: the
 * unfold.  out = {conformal_mean, conformal_max, area_ratio_mean}, cnt = {flips, folds}. */
void orc_chart_distortion(const double* p3, const uint32_t* tris, int64_t ntri, const double* uv, double* out,
                          int64_t* cnt) {
    double conf_sum = 0.0, conf_max = 0.0, area_sum = 0.0;
    int64_t nconf = 0, flips = 0;
    for (int64_t t = 0; t < ntri; ++t) {
        const uint32_t a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        const v3 a3 = P3(p3, a), b3 = P3(p3, b), c3 = P3(p3, c);
        const v2 u0 = UV(uv, a), u1 = UV(uv, b), u2 = UV(uv, c);
        const v3 e1 = v3sub(b3, a3);
        const v3 e2 = v3sub(c3, a3);
        const v3 nrm = v3cross(e1, e2);
        const double ln = v3len(nrm);
        if (ln < 1e-30) continue;
        const v3 n = v3div(nrm, ln);
        const v3 t1 = v3div(e1, fmax(v3len(e1), 1e-30));
        const v3 t2 = v3cross(n, t1);
        const double e1t1 = v3dot(e1, t1), e2t1 = v3dot(e2, t1), e1t2 = v3dot(e1, t2), e2t2 = v3dot(e2, t2);
        const double det = e1t1 * e2t2 - e2t1 * e1t2;
        if (fabs(det) < 1e-30) {
            ++flips;
            continue;
        }
        const double ip00 = e2t2 / det, ip01 = -e2t1 / det, ip10 = -e1t2 / det, ip11 = e1t1 / det;
        const v2 du1 = {u1.u - u0.u, u1.v - u0.v}, du2 = {u2.u - u0.u, u2.v - u0.v};
        const double f[2][2] = {{du1.u * ip00 + du2.u * ip10, du1.u * ip01 + du2.u * ip11},
                                {du1.v * ip00 + du2.v * ip10, du1.v * ip01 + du2.v * ip11}};
        double s1, s2;
        svd2(f, &s1, &s2);
        const double conf = s1 / fmax(s2, 1e-30);
        conf_sum += conf;
        conf_max = fmax(conf_max, conf);
        ++nconf;
        const double uv_area = 0.5 * fabs(du1.u * du2.v - du2.u * du1.v);
        area_sum += uv_area / (0.5 * ln);
    }
    const double den = (double)(nconf > 0 ? nconf : 1);
    out[0] = conf_sum / den;
    out[1] = conf_max;
    out[2] = area_sum / den;
    cnt[0] = flips;
    cnt[1] = orc_winding_flips(uv, tris, ntri);
}

/* rectangularize (metrics.rs:147-227): the rotation of a 72-step ladder over [0, pi] with the smallest
 * sqrt(ext0 (ext0 + 1e-30) + ext1 (ext1 + 1e-30)) of the border's bounding box (first one on ties), the whole
 * chart rotated and normalised into [0,1]^2, border points snapped onto the nearest rectangle side.  q (nv*2)
 * the rectified uv, ext (2) the extent before normalisation.  Without border the chart is only normalised. */
void orc_rectangularize(const double* uv, int64_t nv, const int64_t* bnd, int64_t nb, double* q, double* ext) {
    if (nv == 0) {
        ext[0] = ext[1] = 0.0;
        return;
    }
    if (nb == 0) {
        v2 mn = UV(uv, 0), mx = UV(uv, 0);
        for (int64_t i = 0; i < nv; ++i) {
            const v2 p = UV(uv, i);
            mn.u = fmin(mn.u, p.u); mn.v = fmin(mn.v, p.v);
            mx.u = fmax(mx.u, p.u); mx.v = fmax(mx.v, p.v);
        }
        const v2 e = {mx.u - mn.u, mx.v - mn.v};
        for (int64_t i = 0; i < nv; ++i) {
            const v2 p = UV(uv, i);
            q[2 * i] = (p.u - mn.u) / fmax(e.u, 1e-30);
            q[2 * i + 1] = (p.v - mn.v) / fmax(e.v, 1e-30);
        }
        ext[0] = e.u;
        ext[1] = e.v;
        return;
    }
    const int n_steps = 72;
    int have = 0;
    double bc = 1.0, bs = 0.0, bscore = INFINITY;
    for (int k = 0; k < n_steps; ++k) {
        const double ang = 3.14159265358979323846 * (double)k / (double)(n_steps - 1);
        const double cs = cos(ang), sn = sin(ang);
        double mn = INFINITY, mx = -INFINITY, mnv = INFINITY, mxv = -INFINITY;
        for (int64_t j = 0; j < nb; ++j) {
            const v2 p = UV(uv, (size_t)bnd[j]);
            const double qu = cs * p.u + sn * p.v;
            const double qv = -sn * p.u + cs * p.v;
            mn = fmin(mn, qu); mx = fmax(mx, qu);
            mnv = fmin(mnv, qv); mxv = fmax(mxv, qv);
        }
        const double e0 = mx - mn, e1 = mxv - mnv;
        const double score = sqrt(e0 * (e0 + 1e-30) + e1 * (e1 + 1e-30));
        if (have && score >= bscore) continue;
        have = 1; bc = cs; bs = sn; bscore = score;
    }
    for (int64_t i = 0; i < nv; ++i) {
        const v2 p = UV(uv, i);
        q[2 * i] = bc * p.u + bs * p.v;
        q[2 * i + 1] = -bs * p.u + bc * p.v;
    }
    double mnu = q[0], mnv = q[1], mxu = q[0], mxv = q[1];
    for (int64_t i = 0; i < nv; ++i) {
        mnu = fmin(mnu, q[2 * i]); mnv = fmin(mnv, q[2 * i + 1]);
        mxu = fmax(mxu, q[2 * i]); mxv = fmax(mxv, q[2 * i + 1]);
    }
    const double eu = mxu - mnu, ev = mxv - mnv;
    for (int64_t i = 0; i < nv; ++i) {
        q[2 * i] = (q[2 * i] - mnu) / fmax(eu, 1e-30);
        q[2 * i + 1] = (q[2 * i + 1] - mnv) / fmax(ev, 1e-30);
    }
    for (int64_t j = 0; j < nb; ++j) {
        const int64_t bi = bnd[j];
        const double pu = q[2 * bi], pv = q[2 * bi + 1];
        const double dx = fmin(pu, 1.0 - pu), dy = fmin(pv, 1.0 - pv);
        if (dx <= dy) q[2 * bi] = (pu < 0.5) ? 0.0 : 1.0;
        else q[2 * bi + 1] = (pv < 0.5) ? 0.0 : 1.0;
    }
    ext[0] = eu;
    ext[1] = ev;
}
