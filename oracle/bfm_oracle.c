/*
 * bfm_oracle.c -- CPU restatement of the parallel BFM colour sweep.
 * TEST INFRASTRUCTURE ONLY (see bfm_oracle.h).  Deliberately simple: byte
 * lattices with linear indexing, no bit packing, no Morton order, no atomics.
 */
#include "bfm_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10, written from the round function published by Salmon et al. */

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void philox_words(uint64_t seed, uint32_t c0, uint64_t age, uint32_t tag, uint32_t out[4])
{
    uint32_t ctr[4] = { c0, (uint32_t)age, (uint32_t)(age >> 32), tag };
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    orc_philox4x32_10(ctr, key, out);
}

uint32_t orc_move_word(uint64_t seed, uint64_t age, uint32_t id)
{
    uint32_t w[4];
    philox_words(seed, id, age, 0u, w);
    return w[0];
}

uint32_t orc_int_hash(uint32_t a)
{
    /* six add/xor-shift steps with the published constants */
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}

uint32_t orc_substep_seed(uint64_t seed, uint64_t age, uint32_t substep)
{
    uint32_t w[4];
    philox_words(seed, substep, age, 3u, w);
    return w[0];
}

uint32_t orc_direction_word(const orc_system *s, uint64_t age, uint32_t substep, uint32_t id)
{
    if (s->rng == ORC_RNG_PHILOX)
        return orc_move_word(s->seed, age, id);
    return orc_int_hash(orc_int_hash(id) ^ orc_substep_seed(s->seed, age, substep));
}

void orc_color_order(uint64_t seed, uint64_t age, uint32_t n_colors, uint32_t *order)
{
    for (uint32_t i = 0; i < n_colors; ++i)
        order[i] = i;
    for (uint32_t i = n_colors; i-- > 1;) {
        uint32_t w[4];
        philox_words(seed, i, age, 1u, w);
        uint32_t j = w[0] % (i + 1u);
        uint32_t t = order[i]; order[i] = order[j]; order[j] = t;
    }
}

/* ------------------------------------------------------------------------ */
/* bond set                                                                   */

static int bond_index(int dx, int dy, int dz)
{
    return (dx & 7) | ((dy & 7) << 3) | ((dz & 7) << 6);
}

void orc_standard_bondset(uint8_t forbidden[512])
{
    static const int base[6][3] = { {2,0,0}, {2,1,0}, {2,1,1}, {2,2,1}, {3,0,0}, {3,1,0} };
    static const int perm[6][3] = { {0,1,2}, {0,2,1}, {1,0,2}, {1,2,0}, {2,0,1}, {2,1,0} };
    for (int i = 0; i < 512; ++i)
        forbidden[i] = 1;
    for (int b = 0; b < 6; ++b)
        for (int p = 0; p < 6; ++p)
            for (int s = 0; s < 8; ++s) {
                int v[3];
                for (int k = 0; k < 3; ++k) {
                    int c = base[b][perm[p][k]];
                    v[k] = (s >> k) & 1 ? -c : c;
                }
                forbidden[bond_index(v[0], v[1], v[2])] = 0;
            }
}

/* ------------------------------------------------------------------------ */
/* lattice helpers                                                            */

typedef struct {
    int32_t L[3];
    int32_t M[3];
    size_t  volume;
} geom_t;

static int make_geom(const orc_system *s, geom_t *g)
{
    for (int a = 0; a < 3; ++a) {
        int32_t L = s->box[a];
        if (L < 8 || (L & (L - 1)) != 0)
            return -1;
        g->L[a] = L;
        g->M[a] = L - 1;
    }
    g->volume = (size_t)g->L[0] * (size_t)g->L[1] * (size_t)g->L[2];
    return 0;
}

static size_t site(const geom_t *g, int32_t x, int32_t y, int32_t z)
{
    return ((size_t)(z & g->M[2]) * (size_t)g->L[1] + (size_t)(y & g->M[1])) * (size_t)g->L[0]
           + (size_t)(x & g->M[0]);
}

/* Directions 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z (UpdaterGPUScBFM_AB_Type.cu:1537-1541). */
/* Linear index of the aligned 2x2x2 block holding a site. */
static size_t block_of(const geom_t *g, int32_t x, int32_t y, int32_t z)
{
    const size_t bx = (size_t)((x & g->M[0]) >> 1), by = (size_t)((y & g->M[1]) >> 1),
                 bz = (size_t)((z & g->M[2]) >> 1);
    return bx + (size_t)(g->L[0] / 2) * (by + (size_t)(g->L[1] / 2) * bz);
}

static void dir_delta(int dir, int32_t d[3])
{
    d[0] = d[1] = d[2] = 0;
    d[dir >> 1] = (dir & 1) ? 1 : -1;
}

/* The 4 sites that a cube with lower corner r enters when it moves in dir. */
static void face_sites(const geom_t *g, const int32_t r[3], int dir, size_t out[4])
{
    int a = dir >> 1;
    int b = (a + 1) % 3, c = (a + 2) % 3;
    int32_t q[3];
    q[a] = (dir & 1) ? r[a] + 2 : r[a] - 1;
    int k = 0;
    for (int db = 0; db < 2; ++db)
        for (int dc = 0; dc < 2; ++dc) {
            q[b] = r[b] + db;
            q[c] = r[c] + dc;
            out[k++] = site(g, q[0], q[1], q[2]);
        }
}

/* The 4 sites that a cube with lower corner r leaves when it moves in dir. */
static void back_sites(const geom_t *g, const int32_t r[3], int dir, size_t out[4])
{
    int a = dir >> 1;
    int b = (a + 1) % 3, c = (a + 2) % 3;
    int32_t q[3];
    q[a] = (dir & 1) ? r[a] : r[a] + 1;
    int k = 0;
    for (int db = 0; db < 2; ++db)
        for (int dc = 0; dc < 2; ++dc) {
            q[b] = r[b] + db;
            q[c] = r[c] + dc;
            out[k++] = site(g, q[0], q[1], q[2]);
        }
}

/* The reference's 9-corner front: corners at r_a + 2*sign, r_b-1..r_b+1,
 * r_c-1..r_c+1 (UpdaterGPUScBFM_AB_Type.cu:436-494, the unrolled form). */
static void front9_sites(const geom_t *g, const int32_t r[3], int dir, size_t out[9])
{
    int a = dir >> 1;
    int b = (a + 1) % 3, c = (a + 2) % 3;
    int32_t q[3];
    q[a] = r[a] + ((dir & 1) ? 2 : -2);
    int k = 0;
    for (int db = -1; db <= 1; ++db)
        for (int dc = -1; dc <= 1; ++dc) {
            q[b] = r[b] + db;
            q[c] = r[c] + dc;
            out[k++] = site(g, q[0], q[1], q[2]);
        }
}

static void cube_sites(const geom_t *g, const int32_t r[3], size_t out[8])
{
    int k = 0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
                out[k++] = site(g, r[0] + dx, r[1] + dy, r[2] + dz);
}

/* ------------------------------------------------------------------------ */
/* invariants                                                                 */

int orc_check_system(const orc_system *s, const int32_t *pos, char *msg, int msglen)
{
    geom_t g;
    if (make_geom(s, &g) != 0) {
        if (msg) snprintf(msg, msglen, "box edges must be powers of two >= 8");
        return 1;
    }
    uint8_t *lat = (uint8_t *)calloc(g.volume, 1);
    if (!lat) {
        if (msg) snprintf(msg, msglen, "out of memory");
        return 9;
    }
    int rc = 0;
    for (uint32_t i = 0; i < s->n && rc == 0; ++i) {
        const int32_t *r = pos + 3 * (size_t)i;
        for (int a = 0; a < 3; ++a)
            if (!s->periodic[a] && (r[a] < 0 || r[a] > g.L[a] - 2)) {
                if (msg) snprintf(msg, msglen, "monomer %u outside the walls on axis %d", i, a);
                rc = 2;
            }
        size_t cs[8];
        cube_sites(&g, r, cs);
        for (int k = 0; k < 8 && rc == 0; ++k) {
            if (lat[cs[k]]) {
                if (msg) snprintf(msg, msglen, "monomer %u overlaps another cube", i);
                rc = 3;
            }
            lat[cs[k]] = 1;
        }
    }
    free(lat);
    for (uint32_t i = 0; i < s->n && rc == 0; ++i) {
        for (uint32_t k = s->nbr_off[i]; k < s->nbr_off[i + 1] && rc == 0; ++k) {
            uint32_t j = s->nbr_idx[k];
            if (j >= s->n) {
                if (msg) snprintf(msg, msglen, "monomer %u has neighbour id %u out of range", i, j);
                rc = 4;
                break;
            }
            int32_t d[3];
            for (int a = 0; a < 3; ++a) {
                d[a] = pos[3 * (size_t)j + a] - pos[3 * (size_t)i + a];
                if (s->periodic[a]) {
                    /* minimum image on periodic axes */
                    int32_t m = d[a] & g.M[a];
                    d[a] = m > g.L[a] / 2 ? m - g.L[a] : m;
                }
            }
            if (d[0] < -3 || d[0] > 3 || d[1] < -3 || d[1] > 3 || d[2] < -3 || d[2] > 3) {
                if (msg) snprintf(msg, msglen, "bond %u-%u too long (%d,%d,%d)", i, j, d[0], d[1], d[2]);
                rc = 5;
            } else if (s->forbidden[bond_index(d[0], d[1], d[2])]) {
                if (msg) snprintf(msg, msglen, "bond %u-%u (%d,%d,%d) not in the bond set", i, j, d[0], d[1], d[2]);
                rc = 6;
            } else if (s->color && s->color[i] == s->color[j]) {
                if (msg) snprintf(msg, msglen, "bonded monomers %u and %u share colour %u", i, j, s->color[i]);
                rc = 7;
            }
        }
    }
    if (rc == 0 && msg && msglen > 0)
        msg[0] = 0;
    return rc;
}

/* ------------------------------------------------------------------------ */
/* colouring                                                                  */

/* Partners that count for the colouring: own links as listed, then every
 * monomer that lists i without i listing it (by increasing id).  Returns the
 * CSR of that graph in freshly allocated arrays. */
static int coloring_graph(uint32_t n, const uint32_t *off, const uint32_t *idx, uint32_t **g_off,
                          uint32_t **g_idx)
{
    uint32_t *deg = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    if (!deg)
        return -1;
    for (uint32_t i = 0; i < n; ++i) {
        deg[i] += off[i + 1] - off[i];
        for (uint32_t k = off[i]; k < off[i + 1]; ++k) {
            const uint32_t j = idx[k];
            int back = 0;
            for (uint32_t q = off[j]; q < off[j + 1]; ++q)
                back |= idx[q] == i;
            if (!back)
                deg[j] += 1;
        }
    }
    uint32_t *o = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    uint32_t *put = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    if (!o || !put) {
        free(deg); free(o); free(put);
        return -1;
    }
    o[0] = 0;
    for (uint32_t i = 0; i < n; ++i)
        o[i + 1] = o[i] + deg[i];
    uint32_t *g = (uint32_t *)malloc(sizeof(uint32_t) * (o[n] ? o[n] : 1));
    if (!g) {
        free(deg); free(o); free(put);
        return -1;
    }
    for (uint32_t i = 0; i < n; ++i) {
        put[i] = o[i];
        for (uint32_t k = off[i]; k < off[i + 1]; ++k)
            g[put[i]++] = idx[k];
    }
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = off[i]; k < off[i + 1]; ++k) {
            const uint32_t j = idx[k];
            int back = 0;
            for (uint32_t q = off[j]; q < off[j + 1]; ++q)
                back |= idx[q] == i;
            if (!back)
                g[put[j]++] = i;
        }
    free(deg);
    free(put);
    *g_off = o;
    *g_idx = g;
    return 0;
}

#define ORC_NO_COLOR 0xFFFFFFFFu
#define ORC_COLOR_LIMIT 255u

/* uncoloured partners of v (each listing counts), the first of them in *first */
static uint32_t uncolored_partners(const uint32_t *o, const uint32_t *g, const uint32_t *color, uint32_t v,
                                   uint32_t *first)
{
    uint32_t count = 0;
    *first = ORC_NO_COLOR;
    for (uint32_t k = o[v]; k < o[v + 1]; ++k) {
        if (color[g[k]] != ORC_NO_COLOR)
            continue;
        if (count == 0)
            *first = g[k];
        ++count;
    }
    return count;
}

int orc_graph_coloring(uint32_t n, const uint32_t *off, const uint32_t *idx, int uniform,
                       uint32_t *color, uint32_t *n_colors)
{
    uint32_t *o = NULL, *g = NULL;
    if (coloring_graph(n, off, idx, &o, &g) != 0)
        return -1;
    uint64_t freq[256];
    memset(freq, 0, sizeof freq);
    uint32_t ncol = 0;
    uint32_t *todo = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    if (!todo) {
        free(o); free(g);
        return -1;
    }
    for (uint32_t i = 0; i < n; ++i)
        color[i] = ORC_NO_COLOR;
    int rc = 0;
    for (uint32_t start = 0; start < n && rc == 0; ++start) {
        if (color[start] != ORC_NO_COLOR)
            continue;
        uint32_t top = 0;          /* stack of monomers with uncoloured partners left */
        uint32_t cur = start;
        for (;;) {
            uint32_t first;
            const uint32_t open = uncolored_partners(o, g, color, cur, &first);
            unsigned char seen[256];
            memset(seen, 0, sizeof seen);
            for (uint32_t k = o[cur]; k < o[cur + 1]; ++k)
                if (color[g[k]] != ORC_NO_COLOR)
                    seen[color[g[k]]] = 1;
            uint32_t c = 0;
            while (c < ORC_COLOR_LIMIT && seen[c])
                ++c;
            if (c == ORC_COLOR_LIMIT) {
                rc = -1;
                break;
            }
            color[cur] = c;
            freq[c] += 1;
            if (c + 1 > ncol)
                ncol = c + 1;
            if (open == 0 && top == 0)
                break;
            if (open > 1)
                todo[top++] = cur;
            if (open > 0) {
                cur = first;
                continue;
            }
            /* dead end: back to the newest stacked monomer that still has an
             * uncoloured partner (one with fewer than two leaves the stack) */
            cur = ORC_NO_COLOR;
            while (top > 0) {
                uint32_t f;
                const uint32_t left = uncolored_partners(o, g, color, todo[top - 1], &f);
                if (left < 2)
                    --top;
                if (f != ORC_NO_COLOR) {
                    cur = f;
                    break;
                }
            }
            if (cur == ORC_NO_COLOR)
                break;
        }
    }
    free(todo);
    if (rc == 0 && uniform && ncol > 0) {
        /* ranking of the colours by use, rarest first, ties in colour order;
         * kept sorted by insertion after every change */
        uint32_t rank_color[256];
        uint64_t rank_count[256];
        for (uint32_t c = 0; c < ncol; ++c) {
            uint32_t at = c;
            while (at > 0 && rank_count[at - 1] > freq[c]) {
                rank_color[at] = rank_color[at - 1];
                rank_count[at] = rank_count[at - 1];
                --at;
            }
            rank_color[at] = c;
            rank_count[at] = freq[c];
        }
        const uint64_t target = ((uint64_t)n + ncol - 1) / ncol;
        for (uint32_t v = 0; v < n; ++v) {
            if (freq[color[v]] <= target)
                continue;
            for (uint32_t r = 0; r < ncol && rank_count[r] < target; ++r) {
                const uint32_t nc = rank_color[r];
                int ok = 1;
                for (uint32_t k = o[v]; k < o[v + 1]; ++k)
                    if (color[g[k]] == nc)
                        ok = 0;
                if (!ok)
                    continue;
                const uint32_t oc = color[v];
                freq[oc] -= 1;
                freq[nc] += 1;
                for (uint32_t q = 0; q < ncol; ++q) {
                    if (rank_color[q] == oc)
                        rank_count[q] -= 1;
                    if (rank_color[q] == nc)
                        rank_count[q] += 1;
                }
                /* stable re-sort: adjacent exchanges of out-of-order pairs */
                for (int swapped = 1; swapped;) {
                    swapped = 0;
                    for (uint32_t q = 0; q + 1 < ncol; ++q)
                        if (rank_count[q] > rank_count[q + 1]) {
                            const uint32_t tc = rank_color[q];
                            const uint64_t tn = rank_count[q];
                            rank_color[q] = rank_color[q + 1];
                            rank_count[q] = rank_count[q + 1];
                            rank_color[q + 1] = tc;
                            rank_count[q + 1] = tn;
                            swapped = 1;
                        }
                }
                color[v] = nc;
            }
        }
    }
    free(o);
    free(g);
    *n_colors = ncol;
    return rc;
}

/* ------------------------------------------------------------------------ */
/* sweep                                                                      */

static int walls_ok(const orc_system *s, const geom_t *g, const int32_t r1[3])
{
    for (int a = 0; a < 3; ++a)
        if (!s->periodic[a] && (r1[a] < 0 || r1[a] > g->L[a] - 2))
            return 0;
    return 1;
}

static int bonds_ok(const orc_system *s, const int32_t *pos, uint32_t i, const int32_t r1[3])
{
    for (uint32_t k = s->nbr_off[i]; k < s->nbr_off[i + 1]; ++k) {
        const int32_t *q = pos + 3 * (size_t)s->nbr_idx[k];
        if (s->forbidden[bond_index(q[0] - r1[0], q[1] - r1[1], q[2] - r1[2])])
            return 0;
    }
    return 1;
}

int orc_sweep(const orc_system *s, int32_t *pos, uint64_t age0, uint32_t n_mcs,
              int form, uint64_t *accepted, uint64_t *proposals)
{
    geom_t g;
    if (make_geom(s, &g) != 0 || form < 0 || form > 2 || s->n_colors == 0 ||
        (s->rng != ORC_RNG_INT_HASH && s->rng != ORC_RNG_PHILOX))
        return -1;
    uint64_t n_acc = 0, n_prop = 0;

    /* members of each colour, increasing id */
    uint32_t *cnt = (uint32_t *)calloc(s->n_colors + 1, sizeof(uint32_t));
    uint32_t *members = (uint32_t *)malloc(sizeof(uint32_t) * (s->n ? s->n : 1));
    uint32_t *prop_id = (uint32_t *)malloc(sizeof(uint32_t) * (s->n ? s->n : 1));
    uint8_t *prop_dir = (uint8_t *)malloc(s->n ? s->n : 1);
    uint8_t *ok2 = (uint8_t *)malloc(s->n ? s->n : 1);
    uint8_t *occ = (uint8_t *)calloc(g.volume, 1);   /* form 0/2: 8-site occupancy; form 1: corners */
    uint8_t *tmp = (uint8_t *)calloc(g.volume, 1);   /* form 0: claim counts; form 1: tentative corners */
    uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * s->n_colors);
    /* accepted moves per aligned 2x2x2 block of sites in the current substep */
    uint8_t *touch = (uint8_t *)calloc(g.volume / 8 + 1, 1);
    if (!cnt || !members || !prop_id || !prop_dir || !ok2 || !occ || !tmp || !order || !touch) {
        free(cnt); free(members); free(prop_id); free(prop_dir); free(ok2); free(occ); free(tmp); free(order);
        free(touch);
        return -2;
    }
    for (uint32_t i = 0; i < s->n; ++i) {
        if (s->color[i] >= s->n_colors) {
            free(cnt); free(members); free(prop_id); free(prop_dir); free(ok2); free(occ); free(tmp); free(order);
            return -3;
        }
        cnt[s->color[i] + 1]++;
    }
    for (uint32_t c = 0; c < s->n_colors; ++c)
        cnt[c + 1] += cnt[c];
    {
        uint32_t *fill = (uint32_t *)malloc(sizeof(uint32_t) * s->n_colors);
        memcpy(fill, cnt, sizeof(uint32_t) * s->n_colors);
        for (uint32_t i = 0; i < s->n; ++i)
            members[fill[s->color[i]]++] = i;
        free(fill);
    }

    /* initial lattice */
    for (uint32_t i = 0; i < s->n; ++i) {
        const int32_t *r = pos + 3 * (size_t)i;
        if (form == 1) {
            occ[site(&g, r[0], r[1], r[2])] = 1;
        } else {
            size_t cs[8];
            cube_sites(&g, r, cs);
            for (int k = 0; k < 8; ++k)
                occ[cs[k]] = 1;
        }
    }

    for (uint32_t m = 0; m < n_mcs; ++m) {
        uint64_t age = age0 + m;
        orc_color_order(s->seed, age, s->n_colors, order);
        for (uint32_t sub = 0; sub < s->n_colors; ++sub) {
            uint32_t c = order[sub];
            uint32_t np = 0;
            /* phase 1: propose and check against the state before the substep */
            for (uint32_t t = cnt[c]; t < cnt[c + 1]; ++t) {
                uint32_t i = members[t];
                int dir = (int)(orc_direction_word(s, age, sub, i) % 6u);
                const int32_t *r0 = pos + 3 * (size_t)i;
                int32_t d[3], r1[3];
                dir_delta(dir, d);
                for (int a = 0; a < 3; ++a)
                    r1[a] = r0[a] + d[a];
                if (!walls_ok(s, &g, r1) || !bonds_ok(s, pos, i, r1))
                    continue;
                int blocked = 0;
                if (form == 1) {
                    size_t f[9];
                    front9_sites(&g, r0, dir, f);
                    for (int k = 0; k < 9; ++k)
                        blocked |= occ[f[k]];
                } else {
                    size_t f[4];
                    face_sites(&g, r0, dir, f);
                    for (int k = 0; k < 4; ++k)
                        blocked |= occ[f[k]];
                }
                if (blocked)
                    continue;
                prop_id[np] = i;
                prop_dir[np] = (uint8_t)dir;
                ++np;
            }
            n_prop += np;

            /* phase 2: reject survivors whose new cubes overlap */
            if (form == 0) {
                for (uint32_t p = 0; p < np; ++p) {
                    size_t f[4];
                    face_sites(&g, pos + 3 * (size_t)prop_id[p], prop_dir[p], f);
                    for (int k = 0; k < 4; ++k)
                        tmp[f[k]]++;
                }
                for (uint32_t p = 0; p < np; ++p) {
                    size_t f[4];
                    face_sites(&g, pos + 3 * (size_t)prop_id[p], prop_dir[p], f);
                    ok2[p] = 1;
                    for (int k = 0; k < 4; ++k)
                        if (tmp[f[k]] != 1)
                            ok2[p] = 0;
                }
                for (uint32_t p = 0; p < np; ++p) {
                    size_t f[4];
                    face_sites(&g, pos + 3 * (size_t)prop_id[p], prop_dir[p], f);
                    for (int k = 0; k < 4; ++k)
                        tmp[f[k]] = 0;
                }
            } else if (form == 1) {
                for (uint32_t p = 0; p < np; ++p) {
                    const int32_t *r0 = pos + 3 * (size_t)prop_id[p];
                    int32_t d[3];
                    dir_delta(prop_dir[p], d);
                    tmp[site(&g, r0[0] + d[0], r0[1] + d[1], r0[2] + d[2])] = 1;
                }
                for (uint32_t p = 0; p < np; ++p) {
                    size_t f[9];
                    front9_sites(&g, pos + 3 * (size_t)prop_id[p], prop_dir[p], f);
                    ok2[p] = 1;
                    for (int k = 0; k < 9; ++k)
                        if (tmp[f[k]])
                            ok2[p] = 0;
                }
                for (uint32_t p = 0; p < np; ++p) {
                    const int32_t *r0 = pos + 3 * (size_t)prop_id[p];
                    int32_t d[3];
                    dir_delta(prop_dir[p], d);
                    tmp[site(&g, r0[0] + d[0], r0[1] + d[1], r0[2] + d[2])] = 0;
                }
            } else {
                /* brute force: new cubes A1, B1 overlap iff every coordinate
                 * differs by at most 1 modulo the box */
                for (uint32_t p = 0; p < np; ++p)
                    ok2[p] = 1;
                for (uint32_t p = 0; p < np; ++p) {
                    int32_t a1[3], dp[3];
                    dir_delta(prop_dir[p], dp);
                    for (int a = 0; a < 3; ++a)
                        a1[a] = pos[3 * (size_t)prop_id[p] + a] + dp[a];
                    for (uint32_t q = p + 1; q < np; ++q) {
                        int32_t b1[3], dq[3];
                        dir_delta(prop_dir[q], dq);
                        int overlap = 1;
                        for (int a = 0; a < 3; ++a) {
                            b1[a] = pos[3 * (size_t)prop_id[q] + a] + dq[a];
                            int32_t diff = (b1[a] - a1[a]) & g.M[a];
                            if (!(diff <= 1 || diff >= g.L[a] - 1))
                                overlap = 0;
                        }
                        if (overlap)
                            ok2[p] = ok2[q] = 0;
                    }
                }
            }

            /* Invariant the product's byte-wise lattice update rests on: no two
             * accepted moves of a substep touch the same aligned 2x2x2 block
             * (old or new corner).  Violations return -4. */
            int shared_block = 0;
            for (uint32_t p = 0; p < np; ++p) {
                if (!ok2[p])
                    continue;
                const int32_t *r = pos + 3 * (size_t)prop_id[p];
                int32_t d[3];
                dir_delta(prop_dir[p], d);
                const size_t b0 = block_of(&g, r[0], r[1], r[2]);
                const size_t b1 = block_of(&g, r[0] + d[0], r[1] + d[1], r[2] + d[2]);
                shared_block |= touch[b0]++ != 0;
                if (b1 != b0)
                    shared_block |= touch[b1]++ != 0;
            }
            for (uint32_t p = 0; p < np; ++p) {
                if (!ok2[p])
                    continue;
                const int32_t *r = pos + 3 * (size_t)prop_id[p];
                int32_t d[3];
                dir_delta(prop_dir[p], d);
                touch[block_of(&g, r[0], r[1], r[2])] = 0;
                touch[block_of(&g, r[0] + d[0], r[1] + d[1], r[2] + d[2])] = 0;
            }
            if (shared_block) {
                free(cnt); free(members); free(prop_id); free(prop_dir); free(ok2); free(occ); free(tmp);
                free(order); free(touch);
                return -4;
            }

            /* apply */
            for (uint32_t p = 0; p < np; ++p) {
                if (!ok2[p])
                    continue;
                int32_t *r = pos + 3 * (size_t)prop_id[p];
                int dir = prop_dir[p];
                int32_t d[3];
                dir_delta(dir, d);
                if (form == 1) {
                    occ[site(&g, r[0], r[1], r[2])] = 0;
                    occ[site(&g, r[0] + d[0], r[1] + d[1], r[2] + d[2])] = 1;
                } else {
                    size_t f[4], b[4];
                    face_sites(&g, r, dir, f);
                    back_sites(&g, r, dir, b);
                    for (int k = 0; k < 4; ++k) {
                        occ[b[k]] = 0;
                        occ[f[k]] = 1;
                    }
                }
                r[0] += d[0]; r[1] += d[1]; r[2] += d[2];
                ++n_acc;
            }
        }
    }
    free(cnt); free(members); free(prop_id); free(prop_dir); free(ok2); free(occ); free(tmp); free(order);
    free(touch);
    if (accepted) *accepted = n_acc;
    if (proposals) *proposals = n_prop;
    return 0;
}

int orc_sweep_parallel(const orc_system *s, int32_t *pos, uint64_t age0, uint32_t n_mcs,
                       int threads, uint64_t *accepted, uint64_t *proposals)
{
    geom_t g;
    if (make_geom(s, &g) != 0 || s->n_colors == 0 || threads < 0 ||
        (s->rng != ORC_RNG_INT_HASH && s->rng != ORC_RNG_PHILOX))
        return -1;
    if (threads == 0)
        threads = omp_get_max_threads();
    const size_t n1 = s->n ? s->n : 1;
    uint32_t *cnt = (uint32_t *)calloc(s->n_colors + 1, sizeof(uint32_t));
    uint32_t *members = (uint32_t *)malloc(sizeof(uint32_t) * n1);
    uint8_t *flag = (uint8_t *)calloc(n1, 1);         /* per member slot: 0, or direction + 1 */
    size_t *claim = (size_t *)malloc(sizeof(size_t) * n1);   /* claimed corner site */
    uint8_t *occ = (uint8_t *)calloc(g.volume, 1);
    uint8_t *tmp = (uint8_t *)calloc(g.volume, 1);
    uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * s->n_colors);
    if (!cnt || !members || !flag || !claim || !occ || !tmp || !order) {
        free(cnt); free(members); free(flag); free(claim); free(occ); free(tmp); free(order);
        return -2;
    }
    for (uint32_t i = 0; i < s->n; ++i) {
        if (s->color[i] >= s->n_colors) {
            free(cnt); free(members); free(flag); free(claim); free(occ); free(tmp); free(order);
            return -3;
        }
        cnt[s->color[i] + 1]++;
    }
    for (uint32_t c = 0; c < s->n_colors; ++c)
        cnt[c + 1] += cnt[c];
    {
        uint32_t *fill = (uint32_t *)malloc(sizeof(uint32_t) * s->n_colors);
        memcpy(fill, cnt, sizeof(uint32_t) * s->n_colors);
        for (uint32_t i = 0; i < s->n; ++i)
            members[fill[s->color[i]]++] = i;
        free(fill);
    }
    for (uint32_t i = 0; i < s->n; ++i) {
        const int32_t *r = pos + 3 * (size_t)i;
        occ[site(&g, r[0], r[1], r[2])] = 1;
    }
    unsigned long long n_acc = 0, n_prop = 0;
    for (uint32_t m = 0; m < n_mcs; ++m) {
        const uint64_t age = age0 + m;
        orc_color_order(s->seed, age, s->n_colors, order);
        for (uint32_t sub = 0; sub < s->n_colors; ++sub) {
            const uint32_t c = order[sub];
            const long b = (long)cnt[c], e = (long)cnt[c + 1];
            unsigned long long np = 0, na = 0;
            /* phase 1: propose and check against the state before the substep */
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : np)
            for (long t = b; t < e; ++t) {
                const uint32_t i = members[t];
                const int dir = (int)(orc_direction_word(s, age, sub, i) % 6u);
                const int32_t *r0 = pos + 3 * (size_t)i;
                int32_t d[3], r1[3];
                dir_delta(dir, d);
                for (int a = 0; a < 3; ++a)
                    r1[a] = r0[a] + d[a];
                flag[t] = 0;
                if (!walls_ok(s, &g, r1) || !bonds_ok(s, pos, i, r1))
                    continue;
                size_t f[9];
                front9_sites(&g, r0, dir, f);
                int blocked = 0;
                for (int k = 0; k < 9; ++k)
                    blocked |= occ[f[k]];
                if (blocked)
                    continue;
                flag[t] = (uint8_t)(dir + 1);
                claim[t] = site(&g, r1[0], r1[1], r1[2]);
                ++np;
            }
            /* claims */
#pragma omp parallel for num_threads(threads) schedule(static)
            for (long t = b; t < e; ++t)
                if (flag[t]) {
#pragma omp atomic write
                    tmp[claim[t]] = 1;
                }
            /* phase 2: no other claimed corner in the front */
#pragma omp parallel for num_threads(threads) schedule(static)
            for (long t = b; t < e; ++t) {
                if (!flag[t])
                    continue;
                size_t f[9];
                front9_sites(&g, pos + 3 * (size_t)members[t], flag[t] - 1, f);
                for (int k = 0; k < 9; ++k)
                    if (tmp[f[k]]) {
                        flag[t] |= 0x80u;
                        break;
                    }
            }
            /* apply (accepted moves touch disjoint sites) */
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : na)
            for (long t = b; t < e; ++t) {
                if (!flag[t] || (flag[t] & 0x80u))
                    continue;
                int32_t *r = pos + 3 * (size_t)members[t];
                int32_t d[3];
                dir_delta(flag[t] - 1, d);
                occ[site(&g, r[0], r[1], r[2])] = 0;
                occ[claim[t]] = 1;
                r[0] += d[0]; r[1] += d[1]; r[2] += d[2];
                ++na;
            }
#pragma omp parallel for num_threads(threads) schedule(static)
            for (long t = b; t < e; ++t)
                if (flag[t]) {
#pragma omp atomic write
                    tmp[claim[t]] = 0;
                }
            n_prop += np;
            n_acc += na;
        }
    }
    free(cnt); free(members); free(flag); free(claim); free(occ); free(tmp); free(order);
    if (accepted) *accepted = n_acc;
    if (proposals) *proposals = n_prop;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* random-sequential updater                                                  */

int orc_random_sequential(const orc_system *s, int32_t *pos, uint64_t age0,
                          uint32_t n_mcs, uint64_t *accepted)
{
    geom_t g;
    if (make_geom(s, &g) != 0 || s->n == 0)
        return -1;
    uint8_t *occ = (uint8_t *)calloc(g.volume, 1);
    if (!occ)
        return -2;
    for (uint32_t i = 0; i < s->n; ++i) {
        size_t cs[8];
        cube_sites(&g, pos + 3 * (size_t)i, cs);
        for (int k = 0; k < 8; ++k)
            occ[cs[k]] = 1;
    }
    uint64_t n_acc = 0;
    for (uint32_t m = 0; m < n_mcs; ++m) {
        uint64_t age = age0 + m;
        for (uint32_t k = 0; k < s->n; ++k) {
            uint32_t w[4];
            philox_words(s->seed, k, age, 2u, w);
            uint32_t i = w[0] % s->n;
            int dir = (int)(w[1] % 6u);
            int32_t *r0 = pos + 3 * (size_t)i;
            int32_t d[3], r1[3];
            dir_delta(dir, d);
            for (int a = 0; a < 3; ++a)
                r1[a] = r0[a] + d[a];
            if (!walls_ok(s, &g, r1) || !bonds_ok(s, pos, i, r1))
                continue;
            size_t f[4], b[4];
            face_sites(&g, r0, dir, f);
            if (occ[f[0]] | occ[f[1]] | occ[f[2]] | occ[f[3]])
                continue;
            back_sites(&g, r0, dir, b);
            for (int q = 0; q < 4; ++q) {
                occ[b[q]] = 0;
                occ[f[q]] = 1;
            }
            r0[0] = r1[0]; r0[1] = r1[1]; r0[2] = r1[2];
            ++n_acc;
        }
    }
    free(occ);
    if (accepted) *accepted = n_acc;
    return 0;
}
