/*
 * dendro_oracle.c -- CPU restatement of the Dendro4 octree-FEM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product library never links it and never
 * falls back to it.
 *
 * Every function restates the reference algorithm sequentially and cites the
 * reference file:line it follows (paths relative to the Dendro4 source root).
 * It deliberately works with the reference's own comparison-based logic
 * (TreeNode::operator<, getNext, getFirstChild, lower_bound with less_equal,
 * maxLowerBound) rather than with the 128-bit SFC keys the GPU path uses, so
 * the two implementations are independent.
 *
 * Build: gcc -O2 -fPIC -shared -fopenmp -o oracle/_build/liboracle.so oracle/dendro_oracle.c
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "dendro_oracle.h"

/* ------------------------------------------------------------------------- */
/* SFC tables.  3-D Hilbert: 24 rotations.  rot[r*16 + k], k<8: Morton index */
/* of the Hilbert rank-k child; rot[r*16 + 8 + m]: Hilbert rank of Morton     */
/* child m.  ht[r*8 + m]: rotation inside Morton child m.  Data as in        */
/* src/hcurvedata.cpp:329-546 (HILBERT_ORDERING, pDim==3).                    */
/* ------------------------------------------------------------------------- */
static const char ROT_H[24 * 16 + 1] =
    "0231576403127465" "0451376203741265" "0264573107163425" "6204513725163407"
    "3751046243705261" "5731026443527061" "6237510465127403" "0462375107341625"
    "6402315725341607" "6457310265741203" "0154673201763245" "3201546723105467"
    "6754013245763201" "5764023147563021" "3267540167105423" "5104623721563047"
    "3762045147305621" "6732015445327601" "5401326723541067" "3157640261705243"
    "5137620461527043" "0132675401327645" "5467320167541023" "3102645721305647";

static const unsigned char HT_H[192] = {
    1, 3, 2, 0, 6, 4, 5, 0,   0, 8, 9, 5, 7, 1, 4, 1,   10, 14, 0, 13, 11, 12, 2, 2,
    3, 3, 6, 15, 14, 10, 12, 11,   18, 17, 12, 11, 4, 4, 1, 16,   11, 12, 5, 5, 17, 18, 0, 13,
    20, 6, 3, 6, 0, 19, 9, 2,   21, 22, 18, 17, 1, 16, 7, 7,   8, 8, 22, 21, 9, 23, 17, 18,
    19, 9, 1, 20, 8, 9, 6, 7,   2, 21, 19, 12, 23, 10, 13, 10,   11, 6, 14, 4, 11, 1, 18, 3,
    15, 12, 4, 10, 16, 12, 3, 17,   3, 16, 13, 2, 4, 15, 13, 5,   22, 2, 11, 19, 14, 9, 14, 0,
    15, 20, 15, 3, 19, 13, 2, 23,   8, 13, 5, 23, 16, 7, 16, 4,   23, 17, 13, 17, 5, 21, 8, 12,
    18, 9, 18, 0, 22, 5, 11, 8,   9, 23, 10, 14, 19, 19, 22, 21,   6, 15, 20, 20, 21, 22, 14, 10,
    7, 10, 15, 21, 20, 17, 16, 21,   14, 7, 22, 6, 18, 20, 22, 1,   23, 19, 20, 16, 23, 8, 7, 15};

/* Morton build (src/hcurvedata.cpp:548-563): identity permutation, rotation 0. */
static const char ROT_M[17] = "0123456701234567";
static const unsigned char HT_M[8] = {0, 0, 0, 0, 0, 0, 0, 0};

static int g_hilbert = 1;

void or_set_ordering(int hilbert) { g_hilbert = hilbert ? 1 : 0; }
int or_get_ordering(void) { return g_hilbert; }

static inline int rot_entry(int r, int k) {
    return g_hilbert ? (ROT_H[r * 16 + k] - '0') : (ROT_M[k] - '0');
}
static inline int ht_entry(int r, int m) { return g_hilbert ? HT_H[r * 8 + m] : HT_M[m]; }

/* Copy the tables out (for tests that check their self-consistency). */
void or_get_tables(uint8_t *rot384, uint8_t *ht192) {
    for (int i = 0; i < 384; i++) rot384[i] = (uint8_t)(ROT_H[i] - '0');
    memcpy(ht192, HT_H, 192);
}

/* ------------------------------------------------------------------------- */
/* TreeNode primitives: include/TreeNode.h:39-472, include/TreeNode.tcc.       */
/* ------------------------------------------------------------------------- */
#define LEV(a) ((a)->lev & 31u)
#define FLAG_BOUNDARY 64u
#define FLAG_NODE 128u
#define FLAG_DEP_ELEM 32u

static inline or_tn tn_make(uint32_t x, uint32_t y, uint32_t z, uint32_t lev, uint32_t md) {
    /* TreeNode(1,x,y,z,lev,dim,maxDepth) -- src/TreeNode.cpp:606-620: weight 1 */
    or_tn t;
    t.x = x; t.y = y; t.z = z; t.lev = lev; t.w = 1; t.dim = 3; t.md = md;
    return t;
}

static inline uint32_t bin_length(uint32_t n) { /* src/binUtils.cpp:24-31 */
    uint32_t len = 1;
    while (n > 1) { n >>= 1; len++; }
    return len;
}

static inline uint32_t morton_at(uint32_t x, uint32_t y, uint32_t z, uint32_t bit) {
    return (((z >> bit) & 1u) << 2) | (((y >> bit) & 1u) << 1) | ((x >> bit) & 1u);
}

/* TreeNode::operator== (TreeNode.tcc:92-107) */
static inline int tn_eq(const or_tn *a, const or_tn *b) {
    return a->x == b->x && a->y == b->y && a->z == b->z && LEV(a) == LEV(b);
}

/* TreeNode::operator< (TreeNode.tcc:122-238) */
int or_tn_less(const or_tn *a, const or_tn *b) {
    if (a->x == b->x && a->y == b->y && a->z == b->z) return LEV(a) < LEV(b);
    uint32_t x1 = a->x, x2 = b->x, y1 = a->y, y2 = b->y, z1 = a->z, z2 = b->z;
    if (g_hilbert) {
        uint32_t md = a->md, len;
        if (LEV(a) > LEV(b)) {
            len = 1u << (a->md - LEV(b));
            if (!((x1 < x2 || x1 >= (x2 + len)) || (y1 < y2 || y1 >= (y2 + len)) || (z1 < z2 || z1 >= (z2 + len))))
                return 0;
        } else if (LEV(a) < LEV(b)) {
            len = 1u << (a->md - LEV(a));
            if (!((x2 < x1 || x2 >= (x1 + len)) || (y2 < y1 || y2 >= (y1 + len)) || (z2 < z1 || z2 >= (z1 + len))))
                return 1;
        }
        uint32_t maxDiff = x1 ^ x2;
        if ((y1 ^ y2) > maxDiff) maxDiff = y1 ^ y2;
        if ((z1 ^ z2) > maxDiff) maxDiff = z1 ^ z2;
        uint32_t bl = bin_length(maxDiff);
        uint32_t ncaX = (x1 >> bl) << bl, ncaY = (y1 >> bl) << bl, ncaZ = (z1 >> bl) << bl;
        uint32_t ncaLev = md - bl;
        int rot = 0;
        for (uint32_t i = 0; i < ncaLev; i++) rot = ht_entry(rot, morton_at(ncaX, ncaY, ncaZ, md - i - 1));
        uint32_t bit = md - ncaLev - 1;
        int i1 = morton_at(x1, y1, z1, bit), i2 = morton_at(x2, y2, z2, bit);
        return rot_entry(rot, 8 + i1) < rot_entry(rot, 8 + i2);
    } else {
        uint32_t x = x1 ^ x2, y = y1 ^ y2, z = z1 ^ z2;
        uint32_t maxC = z, yOrx = y;
        if (yOrx < x) { if ((x ^ yOrx) >= yOrx) yOrx = x; }
        if (maxC < yOrx) { if ((maxC ^ yOrx) >= maxC) maxC = yOrx; }
        if (maxC == z) return z1 < z2;
        else if (maxC == y) return y1 < y2;
        else return x1 < x2;
    }
}

static inline int tn_le(const or_tn *a, const or_tn *b) { return or_tn_less(a, b) || tn_eq(a, b); }
static inline int tn_gt(const or_tn *a, const or_tn *b) { return !or_tn_less(a, b) && !tn_eq(a, b); }

static int tn_cmp(const void *pa, const void *pb) {
    const or_tn *a = (const or_tn *)pa, *b = (const or_tn *)pb;
    if (or_tn_less(a, b)) return -1;
    if (or_tn_less(b, a)) return 1;
    return 0;
}

/* TreeNode::isAncestor (TreeNode.tcc:590-610) */
static inline int tn_is_ancestor(const or_tn *a, const or_tn *b) {
    uint32_t la = 1u << (a->md - LEV(a)), lb = 1u << (b->md - LEV(b));
    return LEV(a) < LEV(b) && b->x >= a->x && b->y >= a->y && b->z >= a->z &&
           b->x + lb <= a->x + la && b->y + lb <= a->y + la && b->z + lb <= a->z + la;
}

/* TreeNode::getParent (TreeNode.tcc:275-287) */
static inline or_tn tn_parent(const or_tn *a) {
    uint32_t pl = LEV(a) > 0 ? LEV(a) - 1 : 0, s = a->md - pl;
    return tn_make((a->x >> s) << s, (a->y >> s) << s, (a->z >> s) << s, pl, a->md);
}

/* calculateTreeNodeRotation (TreeNode.tcc:2052-2084) */
static inline int tn_rotation(const or_tn *a) {
    int rot = 0;
    for (uint32_t i = 0; i < LEV(a); i++) rot = ht_entry(rot, morton_at(a->x, a->y, a->z, a->md - i - 1));
    return rot;
}

/* getFirstChild (TreeNode.tcc:552-584) */
static inline or_tn tn_first_child(const or_tn *a) {
    if (!g_hilbert) { or_tn m = *a; m.lev = LEV(a) + 1; m.w = 1; return m; }
    uint32_t len = a->md - LEV(a) - 1;
    int f = rot_entry(tn_rotation(a), 0);
    return tn_make(a->x + ((uint32_t)(f & 1) << len), a->y + ((uint32_t)((f >> 1) & 1) << len),
                   a->z + ((uint32_t)((f >> 2) & 1) << len), LEV(a) + 1, a->md);
}

/* getNext (TreeNode.tcc:460-550) */
static or_tn tn_next(const or_tn *a) {
    or_tn m = *a;
    if (!g_hilbert) {
        uint32_t mask = 1u << (a->md - LEV(a));
        int i;
        for (i = (int)LEV(&m); i >= 0; --i) {
            m.x ^= mask; if (m.x & mask) break;
            m.y ^= mask; if (m.y & mask) break;
            m.z ^= mask; if (m.z & mask) break;
            mask <<= 1;
        }
        m.lev = (uint32_t)i;
        return m;
    }
    for (int i = (int)LEV(&m); i >= 0; --i) {
        if (i == 0) { /* root: sentinel (TreeNode.tcc:483-492) */
            m = tn_make(0, 0, 1u << m.md, 1, a->md + 1);
            break;
        }
        or_tn par = tn_parent(&m);
        int r = tn_rotation(&par);
        uint32_t mid = a->md - LEV(&m);
        int idx = (int)morton_at(m.x, m.y, m.z, mid);
        int ci = rot_entry(r, 8 + idx);
        uint32_t len = par.md - LEV(&par) - 1;
        if (ci < 7) {
            int ni = rot_entry(r, ci + 1);
            m = tn_make(par.x + ((uint32_t)(ni & 1) << len), par.y + ((uint32_t)((ni >> 1) & 1) << len),
                        par.z + ((uint32_t)((ni >> 2) & 1) << len), LEV(&par) + 1, a->md);
            break;
        } else {
            m = par;
        }
    }
    return m;
}

/* getDLD (TreeNode.tcc:427-458) */
static inline or_tn tn_dld(const or_tn *a) {
    uint32_t len = a->md - LEV(a);
    if (!g_hilbert) {
        uint32_t s = 1u << len;
        return tn_make(a->x + s - 1, a->y + s - 1, a->z + s - 1, a->md, a->md);
    }
    int i1 = rot_entry(tn_rotation(a), 7);
    uint32_t bx = i1 & 1, by = (i1 >> 1) & 1, bz = (i1 >> 2) & 1;
    return tn_make(a->x + (bx << len) - bx, a->y + (by << len) - by, a->z + (bz << len) - bz, a->md, a->md);
}

/* getMortonIndex (TreeNode.tcc:60-76) */
static inline uint32_t tn_morton_index(const or_tn *a) {
    uint32_t len = 1u << (a->md - LEV(a)), lp = len << 1;
    return 4 * ((a->z % lp) / len) + 2 * ((a->y % lp) / len) + ((a->x % lp) / len);
}

/* ------------------------------------------------------------------------- */
/* Small dynamic array of TreeNodes.                                          */
/* ------------------------------------------------------------------------- */
typedef struct { or_tn *a; size_t n, cap; } tvec;
static void tv_push(tvec *v, or_tn t) {
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 1024;
        v->a = (or_tn *)realloc(v->a, v->cap * sizeof(or_tn));
    }
    v->a[v->n++] = t;
}

/* omp_par::merge_sort (include/ompUtils.tcc:78-133): the array is cut into */
/* one run per thread, the runs are sorted in parallel and merged pairwise.  */
/* Restated as a recursive task merge sort with the same comparator; equal   */
/* TreeNodes are identical, so its output equals any correct sort's.         */
static void merge_runs(const or_tn *a, size_t na, const or_tn *b, size_t nb, or_tn *out) {
    size_t i = 0, j = 0, k = 0;
    while (i < na && j < nb) out[k++] = (tn_cmp(&b[j], &a[i]) < 0) ? b[j++] : a[i++];
    while (i < na) out[k++] = a[i++];
    while (j < nb) out[k++] = b[j++];
}
static void msort_rec(or_tn *a, or_tn *tmp, size_t n, int depth) {
    if (n < 16384 || depth <= 0) {
        qsort(a, n, sizeof(or_tn), tn_cmp);
        return;
    }
    size_t h = n / 2;
#pragma omp task shared(a, tmp) if (depth > 0)
    msort_rec(a, tmp, h, depth - 1);
    msort_rec(a + h, tmp + h, n - h, depth - 1);
#pragma omp taskwait
    merge_runs(a, h, a + h, n - h, tmp);
    memcpy(a, tmp, n * sizeof(or_tn));
}
static void par_sort(or_tn *a, size_t n) {
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    if (nt <= 1 || n < 32768) {
        qsort(a, n, sizeof(or_tn), tn_cmp);
        return;
    }
    int depth = 0;
    while ((1 << depth) < 2 * nt) depth++;
    or_tn *tmp = (or_tn *)malloc(n * sizeof(or_tn));
#pragma omp parallel
#pragma omp single
    msort_rec(a, tmp, n, depth);
    free(tmp);
}
int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* seq::makeVectorUnique (seqUtils.tcc:106-): sort + drop equal neighbours. */
static void tv_sort_unique(tvec *v) {
    if (v->n < 2) return;
    par_sort(v->a, v->n);
    size_t k = 1;
    for (size_t i = 1; i < v->n; i++)
        if (!tn_eq(&v->a[i], &v->a[k - 1])) v->a[k++] = v->a[i];
    v->n = k;
}

void or_free(void *p) { free(p); }

/* Sort an array of TreeNodes with operator< (par::sampleSort on one rank  */
/* reduces to omp_par::merge_sort, parUtils.tcc:1848-1876).                 */
void or_sort(or_tn *a, size_t n) { par_sort(a, n); }

/* ------------------------------------------------------------------------- */
/* points2Octree on one rank: src/construct.cpp:911-974 + p2oLocal :976-1077  */
/* ------------------------------------------------------------------------- */

/* std::lower_bound(first,last,val,std::less_equal<TreeNode>) */
static size_t lb_less_equal(const or_tn *a, size_t first, size_t last, const or_tn *val) {
    size_t count = last > first ? last - first : 0;
    while (count > 0) {
        size_t step = count / 2, it = first + step;
        if (tn_le(&a[it], val)) { first = it + 1; count -= step + 1; }
        else count = step;
    }
    return first;
}

/* p2oLocal (construct.cpp:976-1077) starting from leaves = {root}. */
/* Returns 0, or -2 when the cursor would be compared with the past-the-end
 * sentinel of getNext (TreeNode.tcc:483-492: a level-1 node with maxDepth+1).
 * The reference's operator< then computes ncaLev = maxDepth - 31 < 0 and
 * loops ~2^32 times; that only happens when a stale cursor reaches the last
 * octant of the domain, and is reported here instead of reproduced. */
static int p2o_local(or_tn *nodes, size_t num_pts, uint32_t maxNumPts, uint32_t maxDepth, tvec *out) {
    or_tn root = tn_make(0, 0, 0, 0, maxDepth);
    or_tn curr = root;
    or_tn last = tn_dld(&root);
    or_tn next = tn_next(&curr);
    size_t curr_pt = 0, next_pt = curr_pt + maxNumPts;
    while (next_pt < num_pts) {
        while (tn_is_ancestor(&curr, &nodes[next_pt - 1]) & tn_is_ancestor(&curr, &nodes[next_pt]) &
               (LEV(&curr) < maxDepth)) {
            curr = tn_first_child(&curr);
            next = tn_next(&curr);
        }
        /* construct.cpp:1022-1032: at level maxDepth-1 the search window is
         * widened to num_pts-1 (the loop never changes curr_node). */
        uint32_t inc = maxNumPts;
        while (LEV(&curr) == maxDepth - 1) {
            inc <<= 1;
            next_pt += inc;
            if (next_pt > num_pts) { next_pt = num_pts - 1; break; }
        }
        if (next.md != maxDepth) return -2;
        size_t found = lb_less_equal(nodes, curr_pt, next_pt, &next) - curr_pt;
        next_pt = curr_pt + found;
        tv_push(out, curr);
        curr = next;
        next = tn_next(&curr);
        if (next_pt > curr_pt) curr_pt = next_pt;
        next_pt = curr_pt + maxNumPts;
    }
    if (curr.md != maxDepth) return -2;
    while (tn_le(&curr, &last)) {
        or_tn d = tn_dld(&curr);
        while (tn_gt(&d, &last) && LEV(&curr) < maxDepth) { curr = tn_first_child(&curr); d = tn_dld(&curr); }
        tv_push(out, curr);
        if (tn_eq(&d, &last)) break;
        curr = tn_next(&curr);
    }
    /* std::list::unique */
    size_t k = out->n ? 1 : 0;
    for (size_t i = 1; i < out->n; i++)
        if (!tn_eq(&out->a[i], &out->a[k - 1])) out->a[k++] = out->a[i];
    out->n = k;
    return 0;
}

/* Quantisation (construct.cpp:929-945) into level-maxDepth TreeNodes. */
void or_quantise(const double *pts, size_t n, uint32_t maxDepth, or_tn *nodes) {
    double scale = (double)(1u << maxDepth);
    for (size_t i = 0; i < n; i++) {
        uint32_t px = (uint32_t)(pts[3 * i] * scale / 1.0);
        uint32_t py = (uint32_t)(pts[3 * i + 1] * scale / 1.0);
        uint32_t pz = (uint32_t)(pts[3 * i + 2] * scale / 1.0);
        nodes[i] = tn_make(px, py, pz, maxDepth, maxDepth);
    }
}

int or_points2octree(const double *pts, size_t n, uint32_t maxDepth, uint32_t maxNumPts,
                     or_tn **out, size_t *nout) {
    tvec res = {0, 0, 0};
    if (maxDepth == 0 || n <= maxNumPts) { /* construct.cpp:916-920, :948-952 */
        tv_push(&res, tn_make(0, 0, 0, 0, maxDepth));
        *out = res.a; *nout = res.n;
        return 0;
    }
    or_tn *nodes = (or_tn *)malloc(n * sizeof(or_tn));
    or_quantise(pts, n, maxDepth, nodes);
    par_sort(nodes, n);  /* omp_par::merge_sort, construct.cpp:955 */
    int rc = p2o_local(nodes, n, maxNumPts, maxDepth, &res);
    free(nodes);
    if (rc) { free(res.a); *out = NULL; *nout = 0; return rc; }
    *out = res.a; *nout = res.n;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* balanceOctree on one rank (balance.cpp:55-64 -> comboRipple :801-951).     */
/* Restated as TreeNode::balanceSubtree (TreeNode.cpp:884-1343) on the root   */
/* block: the block decomposition + ripple of comboRipple yields the same     */
/* minimal 2:1 (face+edge+corner) balanced refinement.  emulate_limit != 0    */
/* reproduces comboRipple's one-rank result for in.size() > maxNum (1e7,      */
/* octUtils.h:477): splitInp[] is never filled, so the output is empty.       */
/* emulate_limit == 1 uses maxNum = 1e7; a larger value is taken as maxNum    */
/* (for tests of both sides of the limit on small inputs).                    */
/* ------------------------------------------------------------------------- */
int or_balance(const or_tn *in, size_t n, int incCorner, int emulate_limit, or_tn **out, size_t *nout) {
    (void)incCorner; /* in 3-D balanceSubtree always adds edge+corner neighbours (dim==3 || incCorner) */
    tvec res = {0, 0, 0};
    if (n < 2) { /* comboRipple: in.size() < 2 returns the input */
        for (size_t i = 0; i < n; i++) tv_push(&res, in[i]);
        *out = res.a; *nout = res.n;
        return 0;
    }
    if (emulate_limit && n > (emulate_limit == 1 ? 10000000u : (size_t)emulate_limit)) {
        *out = NULL; *nout = 0; return 0;
    }
    uint32_t md = in[0].md;
    uint32_t minL = md, maxL = 0;
    for (size_t i = 0; i < n; i++) {
        if (LEV(&in[i]) < minL) minL = LEV(&in[i]);
        if (LEV(&in[i]) > maxL) maxL = LEV(&in[i]);
    }
    if (maxL - minL < 2) { /* TreeNode.cpp:929-947: already balanced */
        for (size_t i = 0; i < n; i++) tv_push(&res, in[i]);
        *out = res.a; *nout = res.n;
        return 0;
    }
    tvec *W = (tvec *)calloc(md + 1, sizeof(tvec)); /* W[l] = workVec at level l */
    for (size_t i = 0; i < n; i++)
        if (LEV(&in[i]) > 0) { or_tn t = in[i]; tv_push(&W[LEV(&t)], t); }
    uint32_t dom = 1u << md;
    for (uint32_t l = md; l >= 2; l--) {
        tv_sort_unique(&W[l]);
        /* remove brothers: keep the first of each parent (TreeNode.cpp:978-993) */
        size_t k = 0;
        for (size_t j = 0; j < W[l].n; j++) {
            if (k > 0) {
                or_tn p0 = tn_parent(&W[l].a[k - 1]), p1 = tn_parent(&W[l].a[j]);
                if (tn_eq(&p0, &p1)) continue;
            }
            W[l].a[k++] = W[l].a[j];
        }
        W[l].n = k;
        for (size_t j = 0; j < k; j++) {
            or_tn p = tn_parent(&W[l].a[j]);
            uint32_t len = 1u << (md - l);
            for (uint32_t c = 0; c < 8; c++) /* the octant and its 7 brothers */
                tv_push(&res, tn_make(p.x + (c & 1) * len, p.y + ((c >> 1) & 1) * len, p.z + ((c >> 2) & 1) * len, l, md));
            uint32_t plen = len << 1;
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        if (!dx && !dy && !dz) continue;
                        int64_t nx = (int64_t)p.x + dx * (int64_t)plen, ny = (int64_t)p.y + dy * (int64_t)plen,
                                nz = (int64_t)p.z + dz * (int64_t)plen;
                        if (nx < 0 || ny < 0 || nz < 0 || nx >= dom || ny >= dom || nz >= dom) continue;
                        tv_push(&W[l - 1], tn_make((uint32_t)nx, (uint32_t)ny, (uint32_t)nz, l - 1, md));
                    }
        }
        free(W[l].a);
        W[l].a = NULL; W[l].n = W[l].cap = 0;
    }
    if (W[1].n) { /* TreeNode.cpp:1284-1301: the 8 children of the root */
        uint32_t len = 1u << (md - 1);
        for (uint32_t c = 0; c < 8; c++)
            tv_push(&res, tn_make((c & 1) * len, ((c >> 1) & 1) * len, ((c >> 2) & 1) * len, 1, md));
    }
    free(W[1].a);
    free(W);
    /* sort, unique, remove ancestors (TreeNode.cpp:1304-1340) */
    tv_sort_unique(&res);
    size_t k = 0;
    for (size_t i = 0; i < res.n; i++) {
        if (i + 1 < res.n && tn_is_ancestor(&res.a[i], &res.a[i + 1])) continue;
        res.a[k++] = res.a[i];
    }
    res.n = k;
    *out = res.a; *nout = res.n;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Reference checkers (src/test/testUtils.C), restated for linear/complete/  */
/* balanced octrees.  Return 1 on success.                                    */
/* ------------------------------------------------------------------------- */
/* The checkers run their independent per-leaf tests on all threads. */
int or_is_sorted_unique(const or_tn *a, size_t n) {
    int ok = 1;
#pragma omp parallel for schedule(static) reduction(&& : ok)
    for (size_t i = 1; i < n; i++) ok = ok && or_tn_less(&a[i - 1], &a[i]);
    return ok;
}
int or_is_linear(const or_tn *a, size_t n) {
    int ok = 1;
#pragma omp parallel for schedule(static) reduction(&& : ok)
    for (size_t i = 1; i < n; i++) ok = ok && !tn_is_ancestor(&a[i - 1], &a[i]);
    return ok;
}
/* complete (testUtils.C:28-80): the first leaf holds the root's first cell, */
/* each leaf's SFC successor (getNext) is an ancestor-or-self of the next     */
/* leaf, and the last leaf ends at the root's last cell (getDLD).             */
int or_is_complete(const or_tn *a, size_t n) {
    if (!n) return 0;
    if (!or_is_sorted_unique(a, n) || !or_is_linear(a, n)) return 0;
    uint32_t md = a[0].md;
    or_tn root = tn_make(0, 0, 0, 0, md), first = root;
    while (LEV(&first) < md) first = tn_first_child(&first);
    if (!(tn_eq(&a[0], &first) || tn_is_ancestor(&a[0], &first))) return 0;
    int ok = 1;
#pragma omp parallel for schedule(static) reduction(&& : ok)
    for (size_t i = 0; i < n - 1; i++) {
        or_tn nx = tn_next(&a[i]);
        ok = ok && (tn_eq(&nx, &a[i + 1]) || tn_is_ancestor(&nx, &a[i + 1]));
    }
    if (!ok) return 0;
    or_tn d0 = tn_dld(&root), d1 = tn_dld(&a[n - 1]);
    return tn_eq(&d0, &d1);
}

/* seq::maxLowerBound (seqUtils.tcc:59-100): greatest a[i] <= key. */
static int max_lower_bound(const or_tn *a, size_t n, const or_tn *key, size_t *idx) {
    size_t lo = 0, hi = n; /* first index with a[i] > key */
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (tn_le(&a[mid], key)) lo = mid + 1; else hi = mid;
    }
    if (lo == 0) { *idx = 0; return 0; }
    *idx = lo - 1;
    return 1;
}

/* 2:1 balance check with corners (testUtils.C isBalanced): every leaf's     */
/* 26 same-size neighbours are covered by leaves at most one level coarser   */
/* and at most one level finer.  Checks coarser side by search at the        */
/* neighbour's anchor cell and finer side by searching the neighbour box's   */
/* cells adjacent to the shared feature is equivalent to checking from the   */
/* finer leaf; since every pair is visited from both sides, checking only     */
/* "neighbour cover is not coarser than level-1" suffices.                    */
static int leaf_is_balanced(const or_tn *a, size_t n, size_t i);
int or_is_balanced(const or_tn *a, size_t n) {
    int ok = 1;
#pragma omp parallel for schedule(dynamic, 4096) reduction(&& : ok)
    for (size_t i = 0; i < n; i++) ok = ok && leaf_is_balanced(a, n, i);
    return ok;
}
/* The same check for the leaves a[idx[k]] only (a sample of a large tree). */
int or_is_balanced_at(const or_tn *a, size_t n, const uint64_t *idx, size_t m) {
    int ok = 1;
#pragma omp parallel for schedule(dynamic, 1024) reduction(&& : ok)
    for (size_t k = 0; k < m; k++) ok = ok && idx[k] < n && leaf_is_balanced(a, n, (size_t)idx[k]);
    return ok;
}
static int leaf_is_balanced(const or_tn *a, size_t n, size_t i) {
    uint32_t md = a[0].md, dom = 1u << md;
    {
        uint32_t l = LEV(&a[i]), len = 1u << (md - l);
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    if (!dx && !dy && !dz) continue;
                    /* the finest cell of the neighbour region touching this leaf */
                    int64_t cx = dx < 0 ? (int64_t)a[i].x - 1 : (dx > 0 ? (int64_t)a[i].x + len : a[i].x);
                    int64_t cy = dy < 0 ? (int64_t)a[i].y - 1 : (dy > 0 ? (int64_t)a[i].y + len : a[i].y);
                    int64_t cz = dz < 0 ? (int64_t)a[i].z - 1 : (dz > 0 ? (int64_t)a[i].z + len : a[i].z);
                    if (cx < 0 || cy < 0 || cz < 0 || cx >= dom || cy >= dom || cz >= dom) continue;
                    or_tn key = tn_make((uint32_t)cx, (uint32_t)cy, (uint32_t)cz, md, md);
                    size_t idx;
                    if (!max_lower_bound(a, n, &key, &idx)) return 0;
                    if (!(tn_is_ancestor(&a[idx], &key) || tn_eq(&a[idx], &key))) return 0;
                    if (LEV(&a[idx]) + 1 < l) return 0;
                }
    }
    return 1;
}

/* ------------------------------------------------------------------------- */
/* ot::DA on one rank.                                                        */
/*   Part1: odaFactory.cpp:150-205 (addBoundaryNodesType1 octFunctions.cpp:  */
/*          681-770, embed to maxDepth+1, re-sort)                            */
/*   Part2: flagNodesType3 (octFunctions.cpp:1556-2168)                       */
/*   Part3: odaFactory.cpp:222-923 (one rank: no ghosts) + buildNodeList     */
/*          (odaBuildNlist.cpp:47-934, 3020-3110)                            */
/* ------------------------------------------------------------------------- */

/* TreeNode::isBoundaryOctant(POSITIVE) (TreeNode.cpp:247-269) */
static inline uint32_t pos_boundary_flags(const or_tn *a) {
    uint32_t len = 1u << (a->md - LEV(a)), blen = (1u << a->md) - len, f = 0;
    if (a->x == blen) f |= 32u;  /* X_POS_BDY */
    if (a->y == blen) f |= 64u;  /* Y_POS_BDY */
    if (a->z == blen) f |= 128u; /* Z_POS_BDY */
    return f;
}

static int bsearch_exact(const or_tn *a, size_t n, const or_tn *key) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (or_tn_less(&a[mid], key)) lo = mid + 1; else hi = mid;
    }
    return lo < n && tn_eq(&a[lo], key);
}

int or_da_build(const or_tn *oct, size_t n, or_da *da) {
    memset(da, 0, sizeof(*da));
    if (n <= 8) return -1; /* blockPartStage1 asserts nodes.size() > 2^dim (blockPart.cpp:583) */
    uint32_t md0 = oct[0].md, md = md0 + 1, D = 1u << md0;
    /* Part1: embed + positive boundary pseudo-octants */
    tvec all = {0, 0, 0}, bdy = {0, 0, 0};
    for (size_t i = 0; i < n; i++) {
        uint32_t d = LEV(&oct[i]), x = oct[i].x, y = oct[i].y, z = oct[i].z, w = oct[i].w;
        uint32_t f = pos_boundary_flags(&oct[i]);
        if (f) {
            or_tn b;
#define PUSH_BDY(X, Y, Z) do { b = tn_make((X), (Y), (Z), (d + 1) | FLAG_BOUNDARY, md); b.w = 10; tv_push(&bdy, b); } while (0)
            if (f & 32u) PUSH_BDY(D, y, z);
            if (f & 64u) PUSH_BDY(x, D, z);
            if (f & 128u) PUSH_BDY(x, y, D);
            if ((f & 96u) == 96u) PUSH_BDY(D, D, z);
            if ((f & 160u) == 160u) PUSH_BDY(D, y, D);
            if ((f & 192u) == 192u) PUSH_BDY(x, D, D);
            if ((f & 224u) == 224u) PUSH_BDY(D, D, D);
#undef PUSH_BDY
        }
        or_tn e = tn_make(x, y, z, d + 1, md);
        e.w = w;
        tv_push(&all, e);
    }
    for (size_t i = 0; i < bdy.n; i++) tv_push(&all, bdy.a[i]);
    free(bdy.a);
    par_sort(all.a, all.n); /* SFC_treeSort (odaFactory.cpp:186) */
    or_tn *in = all.a;
    size_t N = all.n;

    /* Part2: flagNodesType3 -- keys are the parents of -x/-y/-z neighbours */
    for (size_t i = 0; i < N; i++) {
        uint32_t myLev = LEV(&in[i]);
        int hanging = 0;
        if (myLev != 1) {
            uint32_t cn = tn_morton_index(&in[i]);
            uint32_t s = 1u << (md - myLev), X = in[i].x, Y = in[i].y, Z = in[i].z;
            or_tn keys[3];
            int nk = 0;
#define KEY(a, b, c) do { or_tn t_ = tn_make((a), (b), (c), myLev, md); keys[nk++] = tn_parent(&t_); } while (0)
            switch (cn) {
                case 1: if (Y) KEY(X, Y - s, Z); if (Z) KEY(X, Y, Z - s); if (Y && Z) KEY(X, Y - s, Z - s); break;
                case 2: if (X) KEY(X - s, Y, Z); if (Z) KEY(X, Y, Z - s); if (X && Z) KEY(X - s, Y, Z - s); break;
                case 3: if (Z) KEY(X, Y, Z - s); break;
                case 4: if (X) KEY(X - s, Y, Z); if (Y) KEY(X, Y - s, Z); if (X && Y) KEY(X - s, Y - s, Z); break;
                case 5: if (Y) KEY(X, Y - s, Z); break;
                case 6: if (X) KEY(X - s, Y, Z); break;
                default: break; /* 0 and 7 */
            }
#undef KEY
            for (int k = 0; k < nk; k++)
                if (bsearch_exact(in, N, &keys[k])) hanging = 1;
        }
        if (!hanging) in[i].lev |= FLAG_NODE;
    }

    /* Part3 (one rank): localOcts = in, no pre/post ghosts. */
    size_t elemEnd = 0;
    for (size_t i = 0; i < N; i++) if (!(in[i].lev & FLAG_BOUNDARY)) elemEnd++;
    size_t nelem = elemEnd, lbs = N;
    uint32_t *nlist = (uint32_t *)malloc(8 * nelem * sizeof(uint32_t));
    uint8_t *mask = (uint8_t *)malloc(nelem);
    size_t unresolved = 0;
    for (size_t i = 0; i < nelem; i++) {
        uint32_t d = LEV(&in[i]), x = in[i].x, y = in[i].y, z = in[i].z;
        uint32_t ps = md - d + 1;
        uint32_t parX = (x >> ps) << ps, parY = (y >> ps) << ps, parZ = (z >> ps) << ps;
        uint32_t sz = 1u << (md - d);
        for (int j = 0; j < 8; j++) nlist[8 * i + j] = (uint32_t)lbs;
        for (int j = 0; j < 8; j++) {
            uint32_t bx = j & 1, by = (j >> 1) & 1, bz = (j >> 2) & 1;
            or_tn nodeLoc = tn_make(x + bx * sz, y + by * sz, z + bz * sz, md, md);
            or_tn parLoc = tn_make(parX + bx * (sz << 1), parY + by * (sz << 1), parZ + bz * (sz << 1), md, md);
            size_t idx = 0;
            int foundKey;
            if (j == 0) { /* odaBuildNlist.cpp:614-650 */
                if (!(in[i].lev & FLAG_NODE)) {
                    foundKey = max_lower_bound(in, N, &parLoc, &idx);
                    if (foundKey && !(in[idx].x == parLoc.x && in[idx].y == parLoc.y && in[idx].z == parLoc.z &&
                                      (in[idx].lev & FLAG_NODE)))
                        foundKey = 0;
                    nlist[8 * i] = foundKey ? (uint32_t)idx : (uint32_t)i;
                } else {
                    nlist[8 * i] = (uint32_t)i;
                }
                continue;
            }
            /* POS_SEARCH_BLOCK (odaBuildNlist.cpp:531-600) */
            int findGhost = 0;
            or_tn sKey = nodeLoc;
            uint32_t lastLevel = d;
            foundKey = max_lower_bound(in, N, &sKey, &idx);
            if (foundKey && !(tn_is_ancestor(&in[idx], &sKey) || tn_eq(&in[idx], &sKey))) foundKey = 0;
            int doSecondary = 0;
            if (foundKey) {
                if (in[idx].x == sKey.x && in[idx].y == sKey.y && in[idx].z == sKey.z && (in[idx].lev & FLAG_NODE))
                    nlist[8 * i + j] = (uint32_t)idx;
                else
                    doSecondary = 1;
            } else {
                doSecondary = 1; /* own element: treated as a hanging hit (i >= elemBegin) */
            }
            if (doSecondary) { /* POS_SECONDARY_SEARCH_BLOCK (:510-529) */
                sKey = parLoc;
                lastLevel = d - 1;
                foundKey = max_lower_bound(in, N, &sKey, &idx);
                if (foundKey && !(in[idx].x == sKey.x && in[idx].y == sKey.y && in[idx].z == sKey.z &&
                                  (in[idx].lev & FLAG_NODE)))
                    foundKey = 0;
                if (foundKey) nlist[8 * i + j] = (uint32_t)idx;
                else findGhost = 1;
            }
            if (findGhost) { /* :566-600 */
                or_tn newKey = tn_make(sKey.x, sKey.y, sKey.z, lastLevel, md);
                or_tn nkDld = tn_dld(&newKey);
                size_t k = 1;
                findGhost = 0;
                while (idx + k < N) {
                    or_tn pk = tn_parent(&in[idx + k]);
                    if (tn_eq(&pk, &newKey)) {
                        if (in[idx + k].lev & FLAG_NODE) { k++; continue; }
                        findGhost = 1;
                        break;
                    }
                    if (tn_gt(&in[idx + k], &nkDld)) break;
                    k++;
                }
                nlist[8 * i + j] = findGhost ? (uint32_t)(idx + k) : (uint32_t)lbs;
                da->n_find_ghost++;
            }
        }
        /* hanging mask (odaBuildNlist.cpp:840-934) */
        uint8_t m = 0;
        for (int j = 0; j < 8; j++) {
            uint32_t r = nlist[8 * i + j];
            if (r >= lbs) { unresolved++; continue; }
            uint32_t _x = in[r].x, _y = in[r].y, _z = in[r].z, _d = LEV(&in[r]);
            if (!(in[r].lev & FLAG_NODE)) {
                uint32_t s2 = md - _d + 1;
                _x = (_x >> s2) << s2; _y = (_y >> s2) << s2; _z = (_z >> s2) << s2;
            }
            uint32_t bx = j & 1, by = (j >> 1) & 1, bz = (j >> 2) & 1;
            if (j == 0) { if (!(in[i].lev & FLAG_NODE)) m |= 1; }
            else if ((x + bx * sz) != _x || (y + by * sz) != _y || (z + bz * sz) != _z) m |= (uint8_t)(1u << j);
        }
        mask[i] = m;
    }
    /* isNode marking + sizes (odaBuildNlist.cpp:3010-3095) */
    uint8_t *isNode = (uint8_t *)calloc(lbs, 1);
    for (size_t i = 0; i < nelem; i++)
        for (int j = 0; j < 8; j++)
            if (nlist[8 * i + j] < lbs) isNode[nlist[8 * i + j]] = 1;
    uint32_t elemNodes = 0, bndNodes = 0;
    for (size_t i = 0; i < lbs; i++) {
        if (isNode[i]) in[i].lev |= FLAG_NODE;
        if (isNode[i]) { if (i < elemEnd) elemNodes++; else bndNodes++; }
    }
    free(isNode);
    da->octs = in;
    da->n_local = lbs;
    da->n_elem = nelem;
    da->nlist = nlist;
    da->mask = mask;
    da->node_size = elemNodes;
    da->bnd_node_size = bndNodes;
    da->n_unresolved = unresolved;
    /* compact node numbering: NODE octants in local-buffer order (oda.cpp:863-882) */
    da->local_to_node = (uint32_t *)malloc(lbs * sizeof(uint32_t));
    uint32_t cnt = 0;
    for (size_t i = 0; i < lbs; i++) da->local_to_node[i] = (in[i].lev & FLAG_NODE) ? cnt++ : 0xffffffffu;
    da->n_nodes = cnt;
    return unresolved ? 1 : 0;
}

void or_da_free(or_da *da) {
    free(da->octs); free(da->nlist); free(da->mask); free(da->local_to_node);
    memset(da, 0, sizeof(*da));
}

/* Child number + hanging mask -> element type (odaUtils.tcc:12-93 with the */
/* cNumEtype tables of include/oda/cnumEtypes.h).                           */
static const uint8_t ETYPE_MASKS[8][18] = {
    {0, 4, 2, 6, 16, 20, 18, 22, 14, 30, 84, 86, 94, 50, 54, 62, 118, 126},
    {0, 1, 8, 9, 32, 33, 40, 41, 13, 45, 49, 57, 61, 168, 169, 173, 185, 189},
    {0, 8, 1, 9, 64, 72, 65, 73, 11, 75, 200, 201, 203, 81, 89, 91, 217, 219},
    {0, 2, 4, 6, 128, 130, 132, 134, 7, 135, 162, 166, 167, 196, 198, 199, 230, 231},
    {0, 1, 32, 33, 64, 65, 96, 97, 35, 99, 69, 101, 103, 224, 225, 227, 229, 231},
    {0, 2, 128, 130, 16, 18, 144, 146, 138, 154, 19, 147, 155, 208, 210, 218, 211, 219},
    {0, 4, 16, 20, 128, 132, 144, 148, 21, 149, 140, 156, 157, 176, 180, 181, 188, 189},
    {0, 8, 64, 72, 32, 40, 96, 104, 76, 108, 42, 106, 110, 112, 120, 124, 122, 126}};

int or_elem_type(uint32_t cnum, uint32_t mask) {
    if (!mask) return 0;
    for (int t = 0; t < 18; t++) if (ETYPE_MASKS[cnum][t] == mask) return t;
    return -1;
}

void or_get_etype_masks(uint8_t *out144) { memcpy(out144, ETYPE_MASKS, 144); }

/* Jacobian1MatMult element loop (odaJac.C:275-292, 325-366) on compact node */
/* vectors.  L and M are the type-2 stencils [8][18][8][8]                   */
/* (handleType2Stencils.C:678-1000).  lapFac = massFac = 1 unless given.     */
int or_matvec(const or_da *da, const double *L, const double *M, double lapFac, double massFac,
              const double *x, double *y) {
    uint32_t md = da->octs[0].md;
    double hFac = 1.0 / (double)(1u << (md - 1));
    for (size_t i = 0; i < da->n_nodes; i++) y[i] = 0.0;
    for (size_t e = 0; e < da->n_elem; e++) {
        const or_tn *o = &da->octs[e];
        uint32_t lev = LEV(o);
        double h = hFac * (double)(1u << (md - lev));
        double fac1 = lapFac * h / 2.0, fac2 = massFac * h * h * h / 8.0;
        uint32_t idx[8];
        for (int j = 0; j < 8; j++) idx[j] = da->local_to_node[da->nlist[8 * e + j]];
        uint32_t cn = tn_morton_index(o); /* DA::getChildNumber (oda.tcc:1273-1300) */
        int t = or_elem_type(cn, da->mask[e]);
        if (t < 0) return -1;
        const double *Lc = L + ((size_t)cn * 18 + t) * 64, *Mc = M + ((size_t)cn * 18 + t) * 64;
        for (int k = 0; k < 8; k++)
            for (int j = 0; j < 8; j++)
                y[idx[k]] += ((fac1 * Lc[k * 8 + j]) + (fac2 * Mc[k * 8 + j])) * x[idx[j]];
    }
    return 0;
}

/* Element metadata in the form the matvec kernels consume: child number,    */
/* level and hanging mask per element, compact node ids per element.          */
void or_da_element_view(const or_da *da, uint32_t *conn8, uint8_t *cnum, uint8_t *level, uint8_t *mask) {
    for (size_t e = 0; e < da->n_elem; e++) {
        for (int j = 0; j < 8; j++) conn8[8 * e + j] = da->local_to_node[da->nlist[8 * e + j]];
        cnum[e] = (uint8_t)tn_morton_index(&da->octs[e]);
        level[e] = (uint8_t)LEV(&da->octs[e]);
        mask[e] = da->mask[e];
    }
}

/* ------------------------------------------------------------------------- */
/* Compressed element-to-node maps (compressLut = true).                      */
/* ------------------------------------------------------------------------- */

/* DA::buildNodeList compression, odaBuildNlist.cpp:3120-3262 (one rank: all */
/* elements).  Per element i: the sort order (non-hanging: the axis order of  */
/* the coordinate XORs x ^ (x + sz), :3147-3170; hanging: the Morton child    */
/* number, :3172-3185); the 8 nlist entries sorted ascending; the offsets     */
/* i - nl[0] and nl[j] - nl[j-1] as unsigned, remainder offset % 256 into     */
/* rem[8i+j], quotient (unsigned short)(offset >> 8) appended to quot and bit */
/* j of mask[i] set when it is not zero (:3232-3258).  Returns the number of  */
/* quotients; quot may be NULL to count them.                                 */
size_t or_da_compress(const or_da *da, uint8_t *rem, uint8_t *mask, uint8_t *order, uint16_t *quot) {
    size_t nq = 0;
    for (size_t i = 0; i < da->n_elem; i++) {
        const or_tn *e = &da->octs[i];
        uint32_t d = LEV(e), md = e->md, sz = 1u << (md - d);
        uint32_t x = e->x, y = e->y, z = e->z;
        if (!da->mask[i]) {
            uint32_t _x = x ^ (x + sz), _y = y ^ (y + sz), _z = z ^ (z + sz);
            /* DA_FLAGS::SortOrderType: XYZ 0, YXZ 1, XZY 2, ZXY 3, YZX 4, ZYX 5 (odaUtils.h:108-110) */
            if (_x > _y) order[i] = _y > _z ? 5 : (_x > _z ? 4 : 1);
            else order[i] = _x > _z ? 3 : (_y > _z ? 2 : 0);
        } else {
            uint32_t lp = 1u << (md - d + 1);
            order[i] = (uint8_t)(4 * ((z % lp) / sz) + 2 * ((y % lp) / sz) + (x % lp) / sz);
        }
        uint32_t nl[8];
        for (int j = 0; j < 8; j++) nl[j] = da->nlist[8 * i + j];
        for (int a = 1; a < 8; a++) /* ascending */
            for (int b = a; b > 0 && nl[b - 1] > nl[b]; b--) { uint32_t t = nl[b]; nl[b] = nl[b - 1]; nl[b - 1] = t; }
        mask[i] = 0;
        for (int j = 0; j < 8; j++) {
            uint32_t off = j == 0 ? (uint32_t)i - nl[0] : nl[j] - nl[j - 1];
            uint16_t q = (uint16_t)(off >> 8);
            rem[8 * i + j] = (uint8_t)(off % 256);
            if (q) {
                mask[i] |= (uint8_t)(1u << j);
                if (quot) quot[nq] = q;
                nq++;
            }
        }
    }
    return nq;
}

/* DA::getNodeIndices with compressLut (oda.tcc:1304-3950): the sorted list  */
/* from remainders and quotients (quotients consumed in order, :1316-1337),  */
/* then unsorted.  Hanging element (:1341-3875): its 8 vertex points, where   */
/* a hanging vertex j is moved to the parent's corner j (the cNumEtype        */
/* switch, restated in closed form: parent anchor + 2 sz * bits of j), sorted */
/* with operator< at maxDepth; the k-th smallest gets nn[k].  Otherwise the   */
/* fixed pattern of the element's sort order (:3876-3935).                    */
void or_da_decode(const or_da *da, const uint8_t *rem, const uint8_t *mask, const uint8_t *order,
                  const uint16_t *quot, uint32_t *nlist) {
    static const int pat[6][6] = {{1, 2, 3, 4, 5, 6}, {2, 1, 3, 4, 6, 5}, {1, 4, 5, 2, 3, 6},
                                  {4, 1, 5, 2, 6, 3}, {2, 4, 6, 1, 3, 5}, {4, 2, 6, 1, 5, 3}};
    size_t qc = 0;
    for (size_t i = 0; i < da->n_elem; i++) {
        uint32_t nn[8], *nodes = nlist + 8 * i;
        for (int j = 0; j < 8; j++) {
            nn[j] = rem[8 * i + j];
            if (mask[i] & (1u << j)) nn[j] += (uint32_t)quot[qc++] << 8;
            nn[j] = j == 0 ? (uint32_t)i - nn[0] : nn[j] + nn[j - 1];
        }
        uint8_t hn = da->mask[i];
        if (hn) {
            const or_tn *e = &da->octs[i];
            uint32_t d = LEV(e), md = e->md, sz = 1u << (md - d), c = order[i];
            uint32_t px = e->x - (c & 1u) * sz, py = e->y - ((c >> 1) & 1u) * sz, pz = e->z - ((c >> 2) & 1u) * sz;
            or_tn v[8];
            int idx[8];
            for (int j = 0; j < 8; j++) {
                uint32_t s = (hn >> j) & 1u ? 2 * sz : sz;
                uint32_t bx = (hn >> j) & 1u ? px : e->x, by = (hn >> j) & 1u ? py : e->y,
                         bz = (hn >> j) & 1u ? pz : e->z;
                v[j] = tn_make(bx + (j & 1) * s, by + ((j >> 1) & 1) * s, bz + ((j >> 2) & 1) * s, md, md);
                idx[j] = j;
            }
            for (int a = 1; a < 8; a++)
                for (int b = a; b > 0 && or_tn_less(&v[idx[b]], &v[idx[b - 1]]); b--) {
                    int t = idx[b]; idx[b] = idx[b - 1]; idx[b - 1] = t;
                }
            for (int k = 0; k < 8; k++) nodes[idx[k]] = nn[k];
        } else {
            const int *p = pat[order[i]];
            nodes[0] = nn[0];
            nodes[p[0]] = nn[1]; nodes[p[1]] = nn[2]; nodes[p[2]] = nn[3];
            nodes[p[3]] = nn[4]; nodes[p[4]] = nn[5]; nodes[p[5]] = nn[6];
            nodes[7] = nn[7];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Multigrid hierarchy and inter-grid transfers.                              */
/* ------------------------------------------------------------------------- */

/* TreeNode::getChildNumber (TreeNode.tcc:23-60): the position among the     */
/* siblings along the curve (Hilbert rank under HILBERT_ORDERING).           */
static inline uint32_t tn_child_number(const or_tn *a) {
    if (LEV(a) == 0) return 0;
    or_tn par = tn_parent(a);
    uint32_t m = tn_morton_index(a);
    return g_hilbert ? (uint32_t)rot_entry(tn_rotation(&par), 8 + (int)m) : m;
}

/* coarsenOctree, sequential form (src/coarsen.cpp:64-104): a first child    */
/* followed by at least 7 octants before the next first child starts a      */
/* complete sibling group of leaves, which becomes its parent; everything   */
/* else is copied.  Returns -1 when the input does not start with a first   */
/* child (the reference asserts this).                                      */
int or_coarsen(const or_tn *in, size_t n, or_tn **out, size_t *nout) {
    tvec o = {0, 0, 0};
    if (n < 8) {
        for (size_t i = 0; i < n; i++) tv_push(&o, in[i]);
        *out = o.a; *nout = o.n;
        return 0;
    }
    if (tn_child_number(&in[0]) != 0) return -1;
    size_t prev = 0;
    while (prev < n) {
        size_t curr = prev + 1;
        while (curr < n && tn_child_number(&in[curr]) != 0) curr++;
        if (curr >= prev + 8) {
            or_tn p = tn_parent(&in[prev]);
            p.w = in[prev].w;
            tv_push(&o, p);
            for (size_t i = prev + 8; i < curr; i++) tv_push(&o, in[i]);
        } else {
            for (size_t i = prev; i < curr; i++) tv_push(&o, in[i]);
        }
        prev = curr;
    }
    *out = o.a; *nout = o.n;
    return 0;
}

/* Restriction restrictMatVecType1 (src/omg/restrictMatVec.cpp:486-619) with */
/* the touched flags of dummyRestrictMatVecType1 (:410-484) and the macros   */
/* INTERGRID_TRANSFER_LOOP_BLOCK(_DUMMY), ITLB_DUMMY_FINAL_SET_VALUE and      */
/* ITLB_SET_VALUE_SUPPRESSED_DOFS (:71-400).  One rank: every coarse element */
/* is independent and the ALL and WRITABLE loops both run over the elements. */
/* prolongMatVecType1 (src/omg/prolongMatVec.cpp:154-292) applies the same   */
/* entries transposed, so the entries themselves are the result: each        */
/* contribution c[cidx] += Rval * f[fidx] is written out as (row = coarse    */
/* compact node, col = fine compact node, Rval), in the reference's order.   */
/* Tables: rmat1[8][8][18][8][8], rmat2[8][18][8][8] (cCtr-major) and the    */
/* vertex maps map1[8][8][18][8], map2[8][8][8][18][8], map3[7][2][8][18][8], */
/* map4[7][2][8][8][18][8] (Stencils/RmatType*Stencils.inp, vtxMap.inp).      */
/* supf / supc: suppressed fine / coarse compact nodes (NULL: none).          */
/* Fine elements under a coarse element: the reference advances the fine    */
/* loop to the coarse anchor and then takes the next 8 elements as children */
/* 0..7, which is the Morton order of children.  Under Hilbert ordering the  */
/* children are not in that order, so here the group starts at the first     */
/* fine element inside the coarse one and each child's stencil is selected   */
/* by its own Morton child number; under Morton both readings coincide (and  */
/* are checked to).                                                          */
/* Returns 0; -1 on capacity, -2 on a fine/coarse mismatch, -3 on an unknown */
/* element type, -4 on a non-node index.                                     */
static inline int elem_inside_(const or_da *f, size_t e, const or_tn *c) {
    const or_tn *o = &f->octs[e];
    uint32_t sc = 1u << (c->md - LEV(c));
    return o->x >= c->x && o->x < c->x + sc && o->y >= c->y && o->y < c->y + sc && o->z >= c->z &&
           o->z < c->z + sc;
}

int or_restrict_entries(const or_da *daf, const or_da *dac, const double *rmat1, const double *rmat2,
                        const uint16_t *map1, const uint16_t *map2, const uint16_t *map3, const uint16_t *map4,
                        const uint8_t *supf, const uint8_t *supc, uint32_t *rows, uint32_t *cols, double *vals,
                        size_t *nnz) {
    const size_t cap = *nnz;
    size_t cnt = 0;
    uint8_t *dummy = (uint8_t *)calloc(daf->n_local * 16, 1);  /* FineTouchedDummyStatus::flags[16] */
    uint8_t *touched = (uint8_t *)calloc(daf->n_local * 8, 1); /* FineTouchedStatus::flags[8] */
    int rc = 0;
    for (int pass = 0; pass < 2 && !rc; pass++) {
        /* pass 0: INTERGRID_TRANSFER_LOOP_BLOCK_DUMMY; pass 1: INTERGRID_TRANSFER_LOOP_BLOCK */
        size_t fp = 0;
        for (size_t ce = 0; ce < dac->n_elem && !rc; ce++) {
            const or_tn *co = &dac->octs[ce];
            while (fp < daf->n_elem && !elem_inside_(daf, fp, co)) fp++;
            if (fp >= daf->n_elem) { rc = -2; break; }
            if (!g_hilbert && (daf->octs[fp].x != co->x || daf->octs[fp].y != co->y || daf->octs[fp].z != co->z)) {
                rc = -2;
                break;
            }
            uint32_t cNumCoarse = tn_morton_index(co);
            int ctype = or_elem_type(cNumCoarse, dac->mask[ce]);
            if (ctype < 0) { rc = -3; break; }
            const uint32_t *cIdx = &dac->nlist[8 * ce];
            int same = LEV(&daf->octs[fp]) == LEV(co);
            int nf = same ? 1 : 8;
            if (fp + nf > daf->n_elem) { rc = -2; break; }
            for (int k = 0; k < nf && !rc; k++) {
                size_t fe = fp + k;
                if (!elem_inside_(daf, fe, co) || (!same && LEV(&daf->octs[fe]) != LEV(co) + 1)) { rc = -2; break; }
                uint32_t cNumFine = same ? cNumCoarse : tn_morton_index(&daf->octs[fe]);
                if (!g_hilbert && !same && cNumFine != (uint32_t)k) { rc = -2; break; }
                uint8_t fhn = daf->mask[fe];
                const uint32_t *fIdx = &daf->nlist[8 * fe];
                const double *tab = same ? rmat2 + ((size_t)cNumCoarse * 18 + ctype) * 64
                                         : rmat1 + (((size_t)cNumCoarse * 8 + cNumFine) * 18 + ctype) * 64;
                for (int fCtr = 0; fCtr < 8; fCtr++) {
                    if (fhn & (1u << fCtr)) continue;
                    uint32_t fi = fIdx[fCtr];
                    if (pass == 0) {
                        uint32_t thisLev = LEV(&daf->octs[fe]), refLev = LEV(&daf->octs[fi]);
                        uint8_t *d = &dummy[(size_t)fi * 16 + 2 * fCtr];
                        uint8_t type = same ? (thisLev == refLev ? 0 : 2) : (thisLev == refLev ? 1 : 3);
                        uint8_t scal = (thisLev != refLev && thisLev < refLev) ? 1 : 0;
                        d[0] = (uint8_t)(1u | (scal << 1) | (type << 2) | (same ? 0 : (cNumFine << 4)));
                        d[1] = (uint8_t)(cNumCoarse | ((uint32_t)ctype << 3));
                        continue;
                    }
                    uint32_t fn = daf->local_to_node[fi];
                    if (fn == 0xffffffffu) { rc = -4; break; }
                    uint8_t fl = touched[(size_t)fi * 8 + fCtr];
                    for (int cCtr = 0; cCtr < 8; cCtr++) {
                        if (!(fl & (1u << cCtr))) continue;
                        uint32_t cn = dac->local_to_node[cIdx[cCtr]];
                        if (cn == 0xffffffffu) { rc = -4; break; }
                        if ((supf && supf[fn]) || (supc && supc[cn])) continue;
                        if (cnt >= cap) { rc = -1; break; }
                        rows[cnt] = cn;
                        cols[cnt] = fn;
                        vals[cnt] = tab[cCtr * 8 + fCtr];
                        cnt++;
                    }
                }
            }
            fp += nf;
        }
        if (pass == 0 && !rc) {
            /* INTERGRID_TRANSFER_LOOP_BLOCK_DUMMY_FINAL_W / _A (:324-400) */
            uint32_t maxD = daf->octs[0].md - 1, blen = 1u << maxD;
            for (size_t fe = 0; fe < daf->n_elem; fe++) {
                const or_tn *o = &daf->octs[fe];
                uint8_t fhn = daf->mask[fe];
                uint32_t len = 1u << (maxD - (LEV(o) - 1));
                uint32_t bf = 0;
                if (o->x + len == blen) bf |= 32u;
                if (o->y + len == blen) bf |= 64u;
                if (o->z + len == blen) bf |= 128u;
                int nodes[8], nn = 0;
                nodes[nn++] = 0;
                if (bf) {
                    if (bf & 32u) nodes[nn++] = 1;
                    if (bf & 64u) nodes[nn++] = 2;
                    if (bf & 128u) nodes[nn++] = 4;
                    if ((bf & 96u) == 96u) nodes[nn++] = 3;
                    if ((bf & 192u) == 192u) nodes[nn++] = 6;
                    if ((bf & 160u) == 160u) nodes[nn++] = 5;
                    if ((bf & 224u) == 224u) nodes[nn++] = 7;
                }
                for (int q = 0; q < nn; q++) {
                    int nodeNum = nodes[q];
                    if (fhn & (1u << nodeNum)) continue;
                    size_t idx = nodeNum == 0 ? fe : daf->nlist[8 * fe + nodeNum];
                    const uint8_t *d = &dummy[idx * 16];
                    const uint16_t *mp[8];
                    for (int df = 0; df < 8; df++) {
                        mp[df] = NULL;
                        uint8_t f0 = d[2 * df], f1 = d[2 * df + 1];
                        if (!f0) continue;
                        uint32_t st = (f0 & 12u) >> 2, cc = f1 & 7u, ct = (f1 & 248u) >> 3;
                        uint32_t cf = (f0 & 112u) >> 4, sc = (f0 & 2u) >> 1;
                        if (st >= 2 && df == 0) { rc = -2; break; }
                        switch (st) {
                            case 0: mp[df] = map1 + (((size_t)df * 8 + cc) * 18 + ct) * 8; break;
                            case 1: mp[df] = map2 + ((((size_t)df * 8 + cf) * 8 + cc) * 18 + ct) * 8; break;
                            case 2: mp[df] = map3 + (((((size_t)(df - 1) * 2 + sc) * 8 + cc) * 18) + ct) * 8; break;
                            default:
                                mp[df] = map4 + ((((((size_t)(df - 1) * 2 + sc) * 8 + cf) * 8 + cc) * 18) + ct) * 8;
                                break;
                        }
                    }
                    if (rc) break;
                    uint8_t *t = &touched[idx * 8];
                    for (int f = 0; f < 8; f++) {
                        t[f] = 0;
                        if (!mp[f]) continue;
                        for (int c = 0; c < 8; c++) {
                            uint16_t v = mp[f][c];
                            int skip = 0;
                            for (int df = 0; df < f && !skip; df++) {
                                if (!mp[df]) continue;
                                for (int k = 0; k < 8; k++)
                                    if (mp[df][k] == v) { skip = 1; break; }
                            }
                            if (!skip) t[f] |= (uint8_t)(1u << c);
                        }
                    }
                }
                if (rc) break;
            }
        }
    }
    free(dummy);
    free(touched);
    *nnz = cnt;
    return rc;
}
