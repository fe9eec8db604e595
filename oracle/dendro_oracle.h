/*
 * dendro_oracle.h -- CPU restatement of the Dendro4 octree-FEM path.
 * TEST INFRASTRUCTURE ONLY (see dendro_oracle.c).
 */
#ifndef DENDRO_ORACLE_H
#define DENDRO_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same 28-byte layout as ot::TreeNode (include/TreeNode.h:42):
 * x, y, z, level|flags, weight, dim, maxDepth. */
typedef struct {
    uint32_t x, y, z, lev, w, dim, md;
} or_tn;

typedef struct {
    or_tn *octs;            /* local buffer: elements then positive-boundary pseudo-octants, with flags */
    size_t n_local;         /* m_uiLocalBufferSize */
    size_t n_elem;          /* m_uiElementSize */
    uint32_t *nlist;        /* 8 per element, local-buffer indices (m_uiNlist) */
    uint8_t *mask;          /* hanging mask per element (m_ucpLutMasks[2i+1]) */
    uint32_t node_size;     /* m_uiNodeSize */
    uint32_t bnd_node_size; /* m_uiBoundaryNodeSize */
    uint32_t *local_to_node;/* local index -> compact node id (0xffffffff if not a node) */
    uint32_t n_nodes;
    size_t n_unresolved;    /* node-list entries that would need the second-ring exchange */
    size_t n_find_ghost;    /* entries resolved by the forward ghost scan */
} or_da;

void or_set_ordering(int hilbert);
int or_get_ordering(void);
void or_get_tables(uint8_t *rot384, uint8_t *ht192);
void or_get_etype_masks(uint8_t *out144);
int or_tn_less(const or_tn *a, const or_tn *b);
void or_sort(or_tn *a, size_t n);
int or_num_threads(void); /* OpenMP threads the sorts and checkers use */
void or_free(void *p);

void or_quantise(const double *pts, size_t n, uint32_t maxDepth, or_tn *nodes);
int or_points2octree(const double *pts, size_t n, uint32_t maxDepth, uint32_t maxNumPts, or_tn **out, size_t *nout);
int or_balance(const or_tn *in, size_t n, int incCorner, int emulate_limit, or_tn **out, size_t *nout);

int or_is_sorted_unique(const or_tn *a, size_t n);
int or_is_linear(const or_tn *a, size_t n);
int or_is_complete(const or_tn *a, size_t n);
int or_is_balanced(const or_tn *a, size_t n);
int or_is_balanced_at(const or_tn *a, size_t n, const uint64_t *idx, size_t m);

int or_da_build(const or_tn *oct, size_t n, or_da *da);
void or_da_free(or_da *da);
int or_elem_type(uint32_t cnum, uint32_t mask);
int or_matvec(const or_da *da, const double *L, const double *M, double lapFac, double massFac, const double *x,
              double *y);
void or_da_element_view(const or_da *da, uint32_t *conn8, uint8_t *cnum, uint8_t *level, uint8_t *mask);
size_t or_da_compress(const or_da *da, uint8_t *rem, uint8_t *mask, uint8_t *order, uint16_t *quot);
void or_da_decode(const or_da *da, const uint8_t *rem, const uint8_t *mask, const uint8_t *order,
                  const uint16_t *quot, uint32_t *nlist);
int or_coarsen(const or_tn *in, size_t n, or_tn **out, size_t *nout);
int or_restrict_entries(const or_da *daf, const or_da *dac, const double *rmat1, const double *rmat2,
                        const uint16_t *map1, const uint16_t *map2, const uint16_t *map3, const uint16_t *map4,
                        const uint8_t *supf, const uint8_t *supc, uint32_t *rows, uint32_t *cols, double *vals,
                        size_t *nnz);

#ifdef __cplusplus
}
#endif
#endif
