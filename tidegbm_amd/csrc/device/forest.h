// Device-resident ensemble (internal to the HIP library): node tables for the one-thread-per-row
// predict kernel and the packed 8-byte nodes the batched serving kernel stages through LDS.
#pragma once
#include <hip/hip_runtime.h>

#include <initializer_list>
#include <vector>

#include "dev_common.h"

namespace tgbm {

// x = feature | feature2 << 16, y = threshold (categorical node: index of its 4-word bin mask in
// the forest's mask table), z = left, w = right (absolute node ids); flags in a separate byte
// array: bit0 leaf, bits1-2 split type (0 axis, 1 pair, 2/3 categorical), bit3 missing_go_left.
struct FNode {
    int fx, thr, left, right;
};

// Serving node: ids local to the tree (a tree has at most 255 nodes), uint8 threshold.
struct SNode {
    unsigned short f, f2;
    unsigned char thr, left, right, flags;
};
static_assert(sizeof(SNode) == 8, "SNode layout");

constexpr int SERVE_TREES_PER_CHUNK = 32;   // trees staged in LDS at a time by the serving kernel

}  // namespace tgbm

struct tgbm_forest {
    int F = 0, fp = 0, nb = 255, n_trees = 0, n_nodes = 0;
    int max_chunk_nodes = 1;            // most nodes in any SERVE_TREES_PER_CHUNK consecutive trees
    int max_depth = 0;                  // deepest node of any tree (root depth 0)
    bool numeric_only = true;           // every split is a numeric threshold split
    std::vector<int> toff_h;            // host copy of tree_off
    double init = 0.0;
    tgbm::FNode* nodes = nullptr;
    unsigned char* flags = nullptr;
    tgbm::SNode* snodes = nullptr;
    int* tree_off = nullptr;            // [n_trees + 1] first node of each tree
    double* values = nullptr;
    int* roots = nullptr;
    double* lr = nullptr;
    double* edges = nullptr;
    int* n_edges = nullptr;
    unsigned long long* cmask = nullptr;  // [n_cat][4] bin masks of categorical nodes
    int n_cat = 0;
    std::vector<int> cat_off_h;           // [n_trees + 1] first mask of each tree
    void* block = nullptr;              // one device allocation holding every array above
    hipStream_t stream = nullptr;
    // staging for the single-call predict entry points
    double* Xd = nullptr;
    unsigned char* bins = nullptr;
    unsigned char* bins_col_dummy = nullptr;
    double* outd = nullptr;
    int cap = 0;
    ~tgbm_forest() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : {block, (void*)Xd, (void*)bins, (void*)bins_col_dummy, (void*)outd})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void reserve(int n);
};
