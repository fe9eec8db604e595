// dgx_io.cpp -- the reference's on-disk formats for points and octrees
// (src/octFunctions.cpp:319-488, src/treenode2vtk.cpp), host code only: no
// call here touches the GPU, so files can be read and written before (or
// without) dgx_init.
//
//   .pts         u32 N, then 3N doubles x0 y0 z0 x1 ... (readPtsFromFile /
//                writePtsToFile); N = 0 is written as an empty file.
//   data .pts    u32 N, 3N doubles, N doubles of values (readDataPtsFromFile);
//                writeDataPtsToFile writes N and the points only.
//   .ot text     "dim maxDepth\n" "N\n" then "x y z level\n" per octant
//                (readNodesFromFile / writeNodesToFile); nothing for N = 0.
//   raw dump     28-byte ot::TreeNode records back to back
//                (readNodesFromFile_binary / writeNodesToFile_binary).
//   VTK          legacy ASCII unstructured grid of hexahedra with the cell
//                level, weight and rank fields (treeNodesTovtk).
// Octants are ot::TreeNode rows of 7 x u32: x, y, z, level|flags, weight,
// dim, maxDepth.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "dgx_common.h"

namespace {

struct File {
    FILE *f = nullptr;
    File(const char *path, const char *mode) : f(path ? fopen(path, mode) : nullptr) {}
    ~File() {
        if (f) fclose(f);
    }
};

bool fail_open(const char *what, const char *path) {
    dgx_set_error("%s: cannot open '%s': %s", what, path ? path : "(null)", strerror(errno));
    return false;
}

long file_size(FILE *f) {
    if (fseek(f, 0, SEEK_END) != 0) return -1;
    const long n = ftell(f);
    if (fseek(f, 0, SEEK_SET) != 0) return -1;
    return n;
}

// Reads "u32 N, 3N doubles[, N doubles]"; an empty file holds no points.
// The two-call convention of this file: with dst == NULL (or cap too small)
// only *n is set.
int read_pts_file(const char *what, const char *path, bool with_vals, double *pts, double *vals, size_t cap,
                  size_t *n) {
    File in(path, "rb");
    if (!in.f) return fail_open(what, path), -1;
    const long sz = file_size(in.f);
    if (sz < 0) { dgx_set_error("%s: cannot size '%s'", what, path); return -1; }
    uint32_t cnt = 0;
    if (sz > 0 && fread(&cnt, 4, 1, in.f) != 1) { dgx_set_error("%s: '%s' is truncated", what, path); return -1; }
    const uint64_t need = (sz > 0 ? 4 : 0) + (uint64_t)cnt * 8 * (with_vals ? 4 : 3);
    if ((uint64_t)sz < need) {
        dgx_set_error("%s: '%s' holds %ld bytes, %u points need %llu", what, path, sz, cnt, (unsigned long long)need);
        return -1;
    }
    *n = cnt;
    if (!pts || cap < cnt || (with_vals && !vals)) return 0;
    if (cnt && fread(pts, 8, (size_t)3 * cnt, in.f) != (size_t)3 * cnt) {
        dgx_set_error("%s: short read of '%s'", what, path);
        return -1;
    }
    if (with_vals && cnt && fread(vals, 8, cnt, in.f) != cnt) {
        dgx_set_error("%s: short read of '%s'", what, path);
        return -1;
    }
    return 0;
}

int write_pts_file(const char *what, const char *path, const double *pts, size_t n, bool empty_when_zero) {
    if (n && !pts) { dgx_set_error("%s: null points", what); return -1; }
    if (n > 0xffffffffull) { dgx_set_error("%s: %zu points exceed the u32 count of the format", what, n); return -1; }
    File out(path, "wb");
    if (!out.f) return fail_open(what, path), -1;
    if (n == 0 && empty_when_zero) return 0;
    const uint32_t cnt = (uint32_t)n;
    if (fwrite(&cnt, 4, 1, out.f) != 1 || (n && fwrite(pts, 8, 3 * n, out.f) != 3 * n)) {
        dgx_set_error("%s: short write of '%s'", what, path);
        return -1;
    }
    return 0;
}

// 128-bit SFC position of an octant: the curve ranks of its ancestors'
// children from the root down (3 bits per level, the first level in the top
// bits), the octant's level as tie break so that an ancestor precedes its
// descendants (TreeNode::operator<, include/TreeNode.tcc:122-238).
struct SfcKey {
    unsigned __int128 path;
    uint32_t level;
    bool operator<(const SfcKey &o) const { return path != o.path ? path < o.path : level < o.level; }
};
SfcKey sfc_key(const uint32_t *t, const SfcTables &tab) {
    const uint32_t D = t[6], lev = t[3] & 31u;
    unsigned __int128 path = 0;
    uint32_t r = 0;
    for (uint32_t l = 1; l <= lev && l <= D; ++l) {
        const uint32_t b = D - l;
        const uint32_t m = ((t[0] >> b) & 1u) | (((t[1] >> b) & 1u) << 1) | (((t[2] >> b) & 1u) << 2);
        path |= (unsigned __int128)tab.rank[r * 8 + m] << (3 * (31 - l));
        r = tab.crot[r * 8 + m];
    }
    return SfcKey{path, lev};
}

}  // namespace

extern "C" {

int dgx_read_pts(const char *path, double *pts, size_t cap, size_t *n) {
    if (!n) { dgx_set_error("readPtsFromFile: null count"); return -1; }
    return read_pts_file("readPtsFromFile", path, false, pts, nullptr, cap, n);
}

int dgx_write_pts(const char *path, const double *pts, size_t n) {
    return write_pts_file("writePtsToFile", path, pts, n, true);
}

int dgx_read_data_pts(const char *path, double *pts, double *vals, size_t cap, size_t *n) {
    if (!n) { dgx_set_error("readDataPtsFromFile: null count"); return -1; }
    return read_pts_file("readDataPtsFromFile", path, true, pts, vals, cap, n);
}

int dgx_write_data_pts(const char *path, const double *pts, const double *vals, size_t n) {
    (void)vals;  // the reference writer leaves the values out (octFunctions.cpp:409)
    return write_pts_file("writeDataPtsToFile", path, pts, n, false);
}

int dgx_read_nodes(const char *path, uint32_t *tn, size_t cap, size_t *n) {
    if (!n) { dgx_set_error("readNodesFromFile: null count"); return -1; }
    File in(path, "r");
    if (!in.f) return fail_open("readNodesFromFile", path), -1;
    unsigned dim = 0, md = 0, cnt = 0;
    if (fscanf(in.f, "%u %u %u", &dim, &md, &cnt) != 3) {
        dgx_set_error("readNodesFromFile: '%s' lacks the 'dim maxDepth N' header", path);
        return -1;
    }
    *n = cnt;
    if (!tn || cap < cnt) return 0;
    for (unsigned i = 0; i < cnt; ++i) {
        unsigned x, y, z, l;
        if (fscanf(in.f, "%u %u %u %u", &x, &y, &z, &l) != 4) {
            dgx_set_error("readNodesFromFile: '%s' ends after %u of %u octants", path, i, cnt);
            return -1;
        }
        uint32_t *t = tn + (size_t)7 * i;  // TreeNode(x, y, z, level, dim, maxDepth)
        t[0] = x;
        t[1] = dim > 1 ? y : 0u;
        t[2] = dim > 2 ? z : 0u;
        t[3] = l;
        t[4] = 1;
        t[5] = dim;
        t[6] = md;
    }
    return 0;
}

int dgx_write_nodes(const char *path, const uint32_t *tn, size_t n) {
    if (n && !tn) { dgx_set_error("writeNodesToFile: null octants"); return -1; }
    File out(path, "w");
    if (!out.f) return fail_open("writeNodesToFile", path), -1;
    if (!n) return 0;
    const uint32_t dim = tn[5], md = tn[6];
    for (size_t i = 0; i < n; ++i)
        if (tn[7 * i + 5] != dim || tn[7 * i + 6] != md) {
            dgx_set_error("writeNodesToFile: octant %zu has dim/maxDepth %u/%u, the first %u/%u", i, tn[7 * i + 5],
                          tn[7 * i + 6], dim, md);
            return -1;
        }
    fprintf(out.f, "%u %u\n", dim, md);
    fprintf(out.f, "%u\n", (unsigned)n);
    for (size_t i = 0; i < n; ++i) {
        const uint32_t *t = tn + 7 * i;
        fprintf(out.f, "%u %u %u %u\n", t[0], t[1], t[2], t[3] & 31u);
    }
    return ferror(out.f) ? (dgx_set_error("writeNodesToFile: write error on '%s'", path), -1) : 0;
}

int dgx_read_nodes_binary(const char *path, uint32_t *tn, size_t cap, size_t *n) {
    if (!n) { dgx_set_error("readNodesFromFile_binary: null count"); return -1; }
    File in(path, "rb");
    if (!in.f) return fail_open("readNodesFromFile_binary", path), -1;
    const long sz = file_size(in.f);
    if (sz < 0 || sz % 28) {
        dgx_set_error("readNodesFromFile_binary: '%s' is not a whole number of 28-byte octants", path);
        return -1;
    }
    *n = (size_t)sz / 28;
    if (!tn || cap < *n) return 0;
    if (*n && fread(tn, 28, *n, in.f) != *n) {
        dgx_set_error("readNodesFromFile_binary: short read of '%s'", path);
        return -1;
    }
    return 0;
}

int dgx_write_nodes_binary(const char *path, const uint32_t *tn, size_t n) {
    if (n && !tn) { dgx_set_error("writeNodesToFile_binary: null octants"); return -1; }
    File out(path, "wb");
    if (!out.f) return fail_open("writeNodesToFile_binary", path), -1;
    if (n && fwrite(tn, 28, n, out.f) != n) {
        dgx_set_error("writeNodesToFile_binary: short write of '%s'", path);
        return -1;
    }
    return 0;
}

int dgx_write_vtk(uint32_t *tn, size_t n, int rank, const char *name, int hsorted, int curve) {
    if ((n && !tn) || !name) { dgx_set_error("treeNodesTovtk: null argument"); return -1; }
    if (curve != CURVE_HILBERT && curve != CURVE_MORTON) { dgx_set_error("unknown curve %d", curve); return -1; }
    // sort in place by the curve, as the reference does with std::sort
    const SfcTables &tab = host_sfc(curve);
    std::vector<SfcKey> key(n);
    for (size_t i = 0; i < n; ++i) key[i] = sfc_key(tn + 7 * i, tab);
    std::vector<size_t> ord(n);
    std::iota(ord.begin(), ord.end(), size_t(0));
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return key[a] < key[b]; });
    std::vector<uint32_t> sorted(7 * n);
    for (size_t i = 0; i < n; ++i) memcpy(&sorted[7 * i], tn + 7 * ord[i], 28);
    if (n) memcpy(tn, sorted.data(), 28 * n);

    const std::string file = std::string(name) + (curve == CURVE_HILBERT ? "_H_" : "_M_") + std::to_string(rank) + ".vtk";
    File out(file.c_str(), "w");
    if (!out.f) return fail_open("treeNodesTovtk", file.c_str()), -1;
    FILE *f = out.f;
    const int unit_points = 8, num_data_field = hsorted ? 3 : 2;
    fprintf(f, "# vtk DataFile Version 2.0\nDENDRO OCTREES\nASCII\nDATASET UNSTRUCTURED_GRID\n");
    fprintf(f, "POINTS %d float\n", (int)(n * unit_points));
    for (size_t i = 0; i < n; ++i) {
        const uint32_t *t = tn + 7 * i;
        const uint32_t len = 1u << (t[6] - (t[3] & 31u)), x = t[0], y = t[1], z = t[2];
        fprintf(f, "%u %u %u\n%u %u %u\n%u %u %u\n%u %u %u\n", x, y, z, x + len, y, z, x + len, y + len, z, x,
                y + len, z);
        fprintf(f, "%u %u %u\n%u %u %u\n%u %u %u\n%u %u %u\n", x, y, z + len, x + len, y, z + len, x + len, y + len,
                z + len, x, y + len, z + len);
    }
    fprintf(f, "CELLS %zu %d\n", n, (int)(n * unit_points + n));
    for (size_t i = 0; i < n; ++i) {
        fprintf(f, "%d ", unit_points);
        for (int j = 0; j < unit_points; ++j) fprintf(f, "%d ", (int)(i * unit_points + j));
        fprintf(f, "\n");
    }
    fprintf(f, "CELL_TYPES %d\n", (int)n);
    for (size_t i = 0; i < n; ++i) fprintf(f, "12\n");  // VTK_HEXAHEDRON
    // the field count the reference prints (2, or 3 with hsorted) omits the rank array
    fprintf(f, "FIELD OCTREE_DATA %d\n", num_data_field);
    fprintf(f, "cell_level 1 %d int\n", (int)n);
    for (size_t i = 0; i < n; ++i) fprintf(f, "%u ", tn[7 * i + 3] & 31u);
    fprintf(f, "\nweight 1 %d int\n", (int)n);
    for (size_t i = 0; i < n; ++i) fprintf(f, "%u ", tn[7 * i + 4]);
    fprintf(f, "\nmpirank 1 %d int\n", (int)n);
    for (size_t i = 0; i < n; ++i) fprintf(f, "%d ", rank);
    fprintf(f, "\n");
    if (hsorted) {
        fprintf(f, "hilbert_index 1 %d int \n", (int)n);
        for (size_t i = 0; i < n; ++i) fprintf(f, "%d ", (int)(i + 1));
        fprintf(f, "\n");
    }
    return ferror(f) ? (dgx_set_error("treeNodesTovtk: write error on '%s'", file.c_str()), -1) : 0;
}

}  // extern "C"
