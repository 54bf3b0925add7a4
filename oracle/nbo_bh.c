/*
 * nbo_bh.c - oracle: Barnes-Hut build, centre-of-mass pass, walk and scatter.
 *
 * TEST INFRASTRUCTURE ONLY (see nbo.h). Restates src/c/barnes_hut/barnes_hut.c
 * (bounding cube :453-483, Morton keys :285-305, stable LSD radix :312-339,
 * gather :402-440, recursive insertion :115-212 and :493-520, centre-of-mass
 * pass :214-253, statistics :255-283, walk :549-636, driver :638-690,
 * scatter :692-704) with the same FP64 operation order.
 */
#include "nbo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { NBO_MAX_DEPTH = 128, NBO_STACK = 1024 };

typedef struct nbo_node {
    int32_t child[8];
    int32_t leaf; /* sorted particle index, -1 for an internal node */
    int32_t count;
    double com[3];
    double mass;
} nbo_node;

struct nbo_bh {
    double theta;
    double center[3];
    double half;
    size_t n;
    uint64_t* keys;   /* original order */
    size_t* perm;     /* perm[sorted] = original */
    size_t* work;
    double* sx;       /* sorted positions and masses */
    double* sy;
    double* sz;
    double* sm;
    double* sax;
    double* say;
    double* saz;
    nbo_node* nodes;
    size_t node_count;
    size_t node_capacity;
    NBodySimProBarnesHutStats stats;
};

nbo_bh* nbo_bh_create(void) {
    nbo_bh* t = (nbo_bh*)calloc(1, sizeof(nbo_bh));
    if (t != NULL) {
        t->theta = 0.7; /* barnes_hut.c:28 */
        t->half = 1.0;
    }
    return t;
}

static void release_arrays(nbo_bh* t) {
    free(t->keys);
    free(t->perm);
    free(t->work);
    free(t->sx);
    free(t->sy);
    free(t->sz);
    free(t->sm);
    free(t->sax);
    free(t->say);
    free(t->saz);
    t->keys = NULL;
    t->perm = t->work = NULL;
    t->sx = t->sy = t->sz = t->sm = t->sax = t->say = t->saz = NULL;
}

void nbo_bh_destroy(nbo_bh* t) {
    if (t == NULL) {
        return;
    }
    release_arrays(t);
    free(t->nodes);
    free(t);
}

void nbo_bh_set_theta(nbo_bh* t, double theta) {
    if (t != NULL) {
        t->theta = theta;
    }
}

/* barnes_hut.c:286-294 */
static uint64_t axis_code(double v, double c, double h) {
    double u = (v - c) / (2.0 * h) + 0.5;
    if (u < 0.0) {
        u = 0.0;
    } else if (u >= 1.0) {
        u = 1.0 - 1.0e-15;
    }
    return (uint64_t)(u * 2097152.0);
}

/* barnes_hut.c:297-305: x -> bit 3b, y -> 3b+1, z -> 3b+2. */
static uint64_t interleave(uint64_t x, uint64_t y, uint64_t z) {
    uint64_t k = 0;
    for (int b = 0; b < 21; ++b) {
        k |= ((x >> b) & 1u) << (3 * b);
        k |= ((y >> b) & 1u) << (3 * b + 1);
        k |= ((z >> b) & 1u) << (3 * b + 2);
    }
    return k;
}

static int grow_nodes(nbo_bh* t, size_t need) {
    if (need <= t->node_capacity) {
        return 1;
    }
    size_t cap = t->node_capacity == 0 ? 16 : t->node_capacity;
    while (cap < need) {
        cap *= 2;
    }
    nbo_node* p = (nbo_node*)realloc(t->nodes, cap * sizeof(nbo_node));
    if (p == NULL) {
        return 0;
    }
    t->nodes = p;
    t->node_capacity = cap;
    return 1;
}

static size_t new_node(nbo_bh* t) {
    nbo_node* nd = &t->nodes[t->node_count];
    for (int c = 0; c < 8; ++c) {
        nd->child[c] = -1;
    }
    nd->leaf = -1;
    nd->count = 0;
    nd->com[0] = nd->com[1] = nd->com[2] = 0.0;
    nd->mass = 0.0;
    return t->node_count++;
}

/* barnes_hut.c:129-142: strict '>' against the FP64 cell centre. */
static int octant(const nbo_bh* t, size_t p, double cx, double cy, double cz) {
    int o = 0;
    if (t->sx[p] > cx) o |= 1;
    if (t->sy[p] > cy) o |= 2;
    if (t->sz[p] > cz) o |= 4;
    return o;
}

static void set_leaf(nbo_bh* t, size_t node, size_t p) {
    nbo_node* nd = &t->nodes[node];
    nd->leaf = (int32_t)p;
    nd->count = 1;
    nd->mass = t->sm[p];
    nd->com[0] = t->sx[p];
    nd->com[1] = t->sy[p];
    nd->com[2] = t->sz[p];
}

/* barnes_hut.c:145-212 */
static int insert(nbo_bh* t, size_t node, size_t p, double cx, double cy, double cz, double h,
                  int depth) {
    if (depth > NBO_MAX_DEPTH) {
        return N_BODY_SIM_PRO_STATUS_OVERFLOW;
    }
    if (t->nodes[node].leaf != -1) {
        const size_t old = (size_t)t->nodes[node].leaf;
        const int oo = octant(t, old, cx, cy, cz);
        if (!grow_nodes(t, t->node_count + 1)) {
            return N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY;
        }
        t->nodes[node].leaf = -1;
        const size_t moved = new_node(t);
        t->nodes[node].child[oo] = (int32_t)moved;
        set_leaf(t, moved, old);
    }
    const int o = octant(t, p, cx, cy, cz);
    int32_t c = t->nodes[node].child[o];
    if (c == -1) {
        if (!grow_nodes(t, t->node_count + 1)) {
            return N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY;
        }
        c = (int32_t)new_node(t);
        t->nodes[node].child[o] = c;
        set_leaf(t, (size_t)c, p);
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    const double ch = 0.5 * h;
    const double nx = cx + ((o & 1) ? h * 0.5 : -h * 0.5);
    const double ny = cy + ((o & 2) ? h * 0.5 : -h * 0.5);
    const double nz = cz + ((o & 4) ? h * 0.5 : -h * 0.5);
    return insert(t, (size_t)c, p, nx, ny, nz, ch, depth + 1);
}

/* barnes_hut.c:214-253: reverse index order is a post-order; children summed 0..7. */
static void centre_of_mass(nbo_bh* t) {
    for (size_t k = t->node_count; k > 0; --k) {
        nbo_node* nd = &t->nodes[k - 1];
        if (nd->leaf != -1 || nd->count == 0) {
            continue;
        }
        double m = 0.0, x = 0.0, y = 0.0, z = 0.0;
        int cnt = 0;
        for (int c = 0; c < 8; ++c) {
            if (nd->child[c] == -1) {
                continue;
            }
            const nbo_node* ch = &t->nodes[nd->child[c]];
            if (ch->count == 0) {
                continue;
            }
            m += ch->mass;
            x += ch->mass * ch->com[0];
            y += ch->mass * ch->com[1];
            z += ch->mass * ch->com[2];
            cnt += ch->count;
        }
        nd->mass = m;
        nd->count = cnt;
        if (m != 0.0) {
            nd->com[0] = x / m;
            nd->com[1] = y / m;
            nd->com[2] = z / m;
        }
    }
}

static size_t depth_of(const nbo_bh* t, size_t node) {
    size_t best = 0;
    for (int c = 0; c < 8; ++c) {
        const int32_t k = t->nodes[node].child[c];
        if (k != -1) {
            const size_t d = depth_of(t, (size_t)k);
            if (d > best) best = d;
        }
    }
    return best + 1;
}

static int ensure_arrays(nbo_bh* t, size_t n) {
    if (n <= t->n && t->keys != NULL) {
        return 1;
    }
    release_arrays(t);
    t->keys = (uint64_t*)malloc(n * sizeof(uint64_t));
    t->perm = (size_t*)malloc(n * sizeof(size_t));
    t->work = (size_t*)malloc(n * sizeof(size_t));
    t->sx = (double*)malloc(n * sizeof(double));
    t->sy = (double*)malloc(n * sizeof(double));
    t->sz = (double*)malloc(n * sizeof(double));
    t->sm = (double*)malloc(n * sizeof(double));
    t->sax = (double*)malloc(n * sizeof(double));
    t->say = (double*)malloc(n * sizeof(double));
    t->saz = (double*)malloc(n * sizeof(double));
    if (!t->keys || !t->perm || !t->work || !t->sx || !t->sy || !t->sz || !t->sm || !t->sax ||
        !t->say || !t->saz) {
        release_arrays(t);
        t->n = 0;
        return 0;
    }
    t->n = n;
    return 1;
}

/* barnes_hut.c:312-339: three stable passes of 21-bit digits. */
static int radix_sort(nbo_bh* t, size_t n) {
    size_t* count = (size_t*)malloc(((size_t)1 << 21) * sizeof(size_t));
    if (count == NULL) {
        return 0;
    }
    for (int pass = 0; pass < 3; ++pass) {
        const unsigned shift = (unsigned)(21 * pass);
        const uint64_t mask = ((uint64_t)1 << 21) - 1;
        memset(count, 0, ((size_t)1 << 21) * sizeof(size_t));
        for (size_t i = 0; i < n; ++i) {
            ++count[(t->keys[t->perm[i]] >> shift) & mask];
        }
        size_t run = 0;
        for (size_t d = 0; d < ((size_t)1 << 21); ++d) {
            const size_t c = count[d];
            count[d] = run;
            run += c;
        }
        for (size_t i = 0; i < n; ++i) {
            const size_t p = t->perm[i];
            t->work[count[(t->keys[p] >> shift) & mask]++] = p;
        }
        size_t* tmp = t->perm;
        t->perm = t->work;
        t->work = tmp;
    }
    free(count);
    return 1;
}

int nbo_bh_build(nbo_bh* t, const NBodySimProParticleSystemView* view) {
    if (t == NULL || view == NULL) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    const size_t n = view->particle_count;
    if (n == 0) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    /* barnes_hut.c:453-483 */
    double lo[3] = {view->positions_x[0], view->positions_y[0], view->positions_z[0]};
    double hi[3] = {lo[0], lo[1], lo[2]};
    for (size_t i = 1; i < n; ++i) {
        const double p[3] = {view->positions_x[i], view->positions_y[i], view->positions_z[i]};
        for (int a = 0; a < 3; ++a) {
            if (p[a] < lo[a]) lo[a] = p[a];
        }
        for (int a = 0; a < 3; ++a) {
            if (p[a] > hi[a]) hi[a] = p[a];
        }
    }
    double h = 0.5 * (hi[0] - lo[0]);
    const double hy = 0.5 * (hi[1] - lo[1]);
    const double hz = 0.5 * (hi[2] - lo[2]);
    if (hy > h) h = hy;
    if (hz > h) h = hz;
    if (h == 0.0) h = 0.5;
    h *= 1.0 + 1.0e-9;
    for (int a = 0; a < 3; ++a) {
        t->center[a] = 0.5 * (lo[a] + hi[a]);
    }
    t->half = h;

    if (!ensure_arrays(t, n)) {
        return N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY;
    }
    for (size_t i = 0; i < n; ++i) {
        t->perm[i] = i;
        t->keys[i] = interleave(axis_code(view->positions_x[i], t->center[0], h),
                                axis_code(view->positions_y[i], t->center[1], h),
                                axis_code(view->positions_z[i], t->center[2], h));
    }
    if (!radix_sort(t, n)) {
        return N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY;
    }
    for (size_t i = 0; i < n; ++i) {
        const size_t o = t->perm[i];
        t->sx[i] = view->positions_x[o];
        t->sy[i] = view->positions_y[o];
        t->sz[i] = view->positions_z[o];
        t->sm[i] = view->masses[o];
    }

    t->node_count = 0;
    if (!grow_nodes(t, 2 * n - 1)) {
        return N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY;
    }
    new_node(t);
    for (size_t i = 0; i < n; ++i) {
        if (t->nodes[0].count == 0) {
            /* barnes_hut.c:504-512 fills the first root leaf from the caller's
             * (unsorted) arrays at index i; kept as is. */
            nbo_node* r = &t->nodes[0];
            r->leaf = (int32_t)i;
            r->count = 1;
            r->mass = view->masses[i];
            r->com[0] = view->positions_x[i];
            r->com[1] = view->positions_y[i];
            r->com[2] = view->positions_z[i];
            continue;
        }
        const int st = insert(t, 0, i, t->center[0], t->center[1], t->center[2], h, 0);
        if (st != N_BODY_SIM_PRO_STATUS_OK) {
            return st;
        }
    }
    centre_of_mass(t);
    t->stats.node_count = t->node_count;
    t->stats.leaf_count = 0;
    t->stats.internal_node_count = 0;
    for (size_t k = 0; k < t->node_count; ++k) {
        if (t->nodes[k].leaf != -1) {
            ++t->stats.leaf_count;
        } else if (t->nodes[k].count > 0) {
            ++t->stats.internal_node_count;
        }
    }
    t->stats.maximum_depth = depth_of(t, 0);
    t->stats.accepted_approximations = 0;
    t->stats.exact_interactions = 0;
    return N_BODY_SIM_PRO_STATUS_OK;
}

typedef struct stack_entry {
    int32_t node;
    double c[3];
    double h;
} stack_entry;

/* barnes_hut.c:549-636: LIFO walk, children pushed 0..7 (child 7 popped first). */
static void walk_one(const nbo_bh* t, const NBodySimProGravity* g, size_t q, double* out,
                     size_t* approx, size_t* exact) {
    stack_entry st[NBO_STACK];
    int top = 1;
    st[0].node = 0;
    st[0].c[0] = t->center[0];
    st[0].c[1] = t->center[1];
    st[0].c[2] = t->center[2];
    st[0].h = t->half;
    const double qx = t->sx[q], qy = t->sy[q], qz = t->sz[q];
    const double th2 = t->theta * t->theta;
    const double eps2 = g->softening_squared;
    const double G = g->gravitational_constant;
    double ax = out[0], ay = out[1], az = out[2];
    while (top > 0) {
        const stack_entry e = st[--top];
        const nbo_node* nd = &t->nodes[e.node];
        if (nd->count == 0) {
            continue;
        }
        if (nd->leaf != -1) {
            if ((size_t)nd->leaf != q) {
                const size_t j = (size_t)nd->leaf;
                const double dx = t->sx[j] - qx;
                const double dy = t->sy[j] - qy;
                const double dz = t->sz[j] - qz;
                const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                const double inv3 = 1.0 / (r2 * sqrt(r2));
                const double s = G * t->sm[j] * inv3;
                ax += s * dx;
                ay += s * dy;
                az += s * dz;
                ++(*exact);
            }
            continue;
        }
        const double dx = nd->com[0] - qx;
        const double dy = nd->com[1] - qy;
        const double dz = nd->com[2] - qz;
        const double d2 = dx * dx + dy * dy + dz * dz;
        const double size2 = 4.0 * e.h * e.h;
        if (d2 > 0.0 && size2 < th2 * d2) {
            const double r2 = d2 + eps2;
            const double inv3 = 1.0 / (r2 * sqrt(r2));
            const double s = G * nd->mass * inv3;
            ax += s * dx;
            ay += s * dy;
            az += s * dz;
            ++(*approx);
            continue;
        }
        const double ch = 0.5 * e.h;
        const double off = e.h * 0.5;
        for (int c = 0; c < 8; ++c) {
            const int32_t k = nd->child[c];
            if (k == -1) {
                continue;
            }
            if (top >= NBO_STACK) {
                break;
            }
            st[top].node = k;
            st[top].c[0] = e.c[0] + ((c & 1) ? off : -off);
            st[top].c[1] = e.c[1] + ((c & 2) ? off : -off);
            st[top].c[2] = e.c[2] + ((c & 4) ? off : -off);
            st[top].h = ch;
            ++top;
        }
    }
    out[0] = ax;
    out[1] = ay;
    out[2] = az;
}

/* barnes_hut.c:638-704 */
int nbo_bh_compute(nbo_bh* t, const NBodySimProParticleSystemView* view,
                   const NBodySimProGravity* gravity) {
    if (t == NULL || view == NULL || gravity == NULL) {
        return N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT;
    }
    const int st = nbo_bh_build(t, view);
    if (st != N_BODY_SIM_PRO_STATUS_OK) {
        return st;
    }
    const long long n = (long long)view->particle_count;
    size_t total_approx = 0, total_exact = 0;
#pragma omp parallel for schedule(static) reduction(+ : total_approx, total_exact)
    for (long long i = 0; i < n; ++i) {
        size_t a = 0, e = 0;
        double acc[3] = {0.0, 0.0, 0.0};
        walk_one(t, gravity, (size_t)i, acc, &a, &e);
        t->sax[i] = acc[0];
        t->say[i] = acc[1];
        t->saz[i] = acc[2];
        total_approx += a;
        total_exact += e;
    }
    for (long long i = 0; i < n; ++i) {
        const size_t o = t->perm[i];
        view->accelerations_x[o] = t->sax[i];
        view->accelerations_y[o] = t->say[i];
        view->accelerations_z[o] = t->saz[i];
    }
    t->stats.accepted_approximations = total_approx;
    t->stats.exact_interactions = total_exact;
    return N_BODY_SIM_PRO_STATUS_OK;
}

void nbo_bh_stats(const nbo_bh* t, NBodySimProBarnesHutStats* s) {
    if (t != NULL && s != NULL) {
        *s = t->stats;
    }
}

void nbo_bh_root(const nbo_bh* t, double out[4]) {
    out[0] = t->center[0];
    out[1] = t->center[1];
    out[2] = t->center[2];
    out[3] = t->half;
}

void nbo_bh_keys(const nbo_bh* t, uint64_t* keys_out, int64_t* perm_out) {
    for (size_t i = 0; i < t->n; ++i) {
        if (keys_out != NULL) keys_out[i] = t->keys[i];
        if (perm_out != NULL) perm_out[i] = (int64_t)t->perm[i];
    }
}

size_t nbo_bh_node_count(const nbo_bh* t) {
    return t == NULL ? 0 : t->node_count;
}

static void export_rec(const nbo_bh* t, size_t node, int depth, double cx, double cy, double cz,
                       double h, double* rows, size_t* row) {
    const nbo_node* nd = &t->nodes[node];
    double* r = rows + 11 * (*row);
    ++(*row);
    r[0] = depth;
    r[1] = cx;
    r[2] = cy;
    r[3] = cz;
    r[4] = h;
    r[5] = nd->count;
    r[6] = nd->mass;
    r[7] = nd->com[0];
    r[8] = nd->com[1];
    r[9] = nd->com[2];
    r[10] = nd->leaf;
    const double off = h * 0.5;
    for (int c = 0; c < 8; ++c) {
        if (nd->child[c] != -1) {
            export_rec(t, (size_t)nd->child[c], depth + 1, cx + ((c & 1) ? off : -off),
                       cy + ((c & 2) ? off : -off), cz + ((c & 4) ? off : -off), 0.5 * h, rows,
                       row);
        }
    }
}

void nbo_bh_export_nodes(const nbo_bh* t, double* rows) {
    size_t row = 0;
    if (t != NULL && t->node_count > 0) {
        export_rec(t, 0, 0, t->center[0], t->center[1], t->center[2], t->half, rows, &row);
    }
}
