// Hierarchical (H-) matrix structure: a tree whose leaves are Dense or
// LowRank blocks (docs/README.md:15-19 of the reference).
//
// The reference's Hierarchical class is an empty stub in the snapshot
// (matrix/Hierarchical.h:4-7); its API is implied by src/launcher.cu:126-139:
// getNx_abs(), getNy_abs(), getRootIndex() -> h_index*, and
// generateOps_GETRF(h_index*, dev_temp*) -> h_ops_tree*.  Those methods are
// provided here, plus the structure builders and the storage plan.
//
// Storage: every block lives in one arena (pointer-table slot `ptr_id`):
//   * Dense    ny x nx row-major, ld = nx, at element offset `off`;
//   * LowRank  A = U * V^T with U ny x r and V nx x r, both row-major with
//              ld = r, at `off_u` / `off_v`.
// Offsets are element counts aligned to 32 elements (256 bytes in FP64).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

enum class element_t : int { empty = 0, dense = 1, low_rank = 2, hierarchical = 3 };

struct h_block {
  element_t type = element_t::empty;
  int ny = 0, nx = 0;      // rows, columns
  int row0 = 0, col0 = 0;  // absolute position in the root matrix
  int level = 0;           // depth in the tree (root = 0)
  // dense
  int64_t off = 0;
  // low rank
  int rank = 0;
  int64_t off_u = 0, off_v = 0;
  // hierarchical: py x px children, row-major
  int py = 0, px = 0;
  std::vector<int> row_sizes, col_sizes;
  std::vector<h_block*> children;

  ~h_block();
  h_block* child(int i, int j) const { return children[(size_t)i * px + j]; }
};

// Position of a node in the hierarchy (src/launcher.cu:138).
class h_index {
public:
  const h_block* node;
  explicit h_index(const h_block* n) : node(n) {}
};

class dev_temp;
class h_ops_tree;

class Hierarchical {
private:
  h_block* root;
  int64_t storage;  // arena size in elements
  int ptr_id;

  int64_t plan(h_block* b, int64_t cursor);

public:
  // Takes ownership of `root`; computes the storage plan.
  Hierarchical(h_block* root, int ptr_id = 0);
  ~Hierarchical();
  Hierarchical(const Hierarchical&) = delete;
  Hierarchical& operator=(const Hierarchical&) = delete;

  // All-dense binary partition down to `leaf` rows/cols (leaf-size grid).
  static Hierarchical* denseGrid(int n, int leaf, int ptr_id = 0);
  // HODLR: `levels` binary splits of an n x n matrix; off-diagonal blocks of
  // every level are rank-`rank` U*V^T, diagonal leaves are dense.
  static Hierarchical* hodlr(int n, int levels, int rank, int ptr_id = 0);

  int getNx_abs() const { return root->nx; }
  int getNy_abs() const { return root->ny; }
  const h_block* getRoot() const { return root; }
  int64_t getStorage() const { return storage; }
  int getPtrId() const { return ptr_id; }
  void setPtrId(int id) { ptr_id = id; }

  // Caller deletes the returned index (src/launcher.cu:143).
  const h_index* getRootIndex() const { return new h_index(root); }

  // Flattened h-ops of the in-place LU factorization (no pivoting, unit L).
  // `slot` overrides the matrix's pointer-table slot (batched programs).
  h_ops_tree* generateOps_GETRF(const h_index* self, dev_temp* tmp, int slot = -1) const;
  // Flattened h-ops of the forward + backward substitution applied to the
  // nrhs-column row-major block at (rhs_ptr_id, rhs_off, ld = rhs_ld);
  // rhs_split > 1 cuts the columns into independent groups.
  // inv_base >= 0 applies each dense diagonal leaf through the explicit
  // inverses that generateOps_Invert leaves at inv_base of the matrix slot.
  h_ops_tree* generateOps_Solve(const h_index* self, int rhs_ptr_id, int64_t rhs_off, int nrhs, int rhs_ld,
                                int rhs_split = 1, int slot = -1, int64_t inv_base = -1) const;

  // Explicit inverses of the factored dense diagonal leaves.  Leaf i of
  // inverseLayout() keeps L^{-1} (unit lower factor) at inv_base + inv_off
  // and U^{-1} at inv_base + inv_off + align32(nb * nb), row-major, ld = nb.
  struct inv_block {
    int64_t leaf_off;  // arena offset of the leaf
    int nb;            // leaf order
    int64_t inv_off;   // offset of L^{-1} within the inverse area
    int64_t row0;      // first row of the leaf
  };
  void inverseLayout(std::vector<inv_block>& out) const;
  // Elements of the inverse area (leaf inverses, then, for HODLR-shaped
  // trees, the low-rank factors of inverseNodes()).
  int64_t inverseStorage() const;

  // Hierarchical inverse of the factors of a HODLR-shaped tree (every
  // hierarchical diagonal node is 2 x 2 with low-rank off-diagonal children).
  // For a node with diagonal children A, B, the unit-lower factor is
  // [A 0; U_l V_l^T B] and the upper factor [A U_c V_c^T; 0 B], so
  //   L^{-1} has the off-diagonal block -l_u l_v^T, rows of B x columns of A,
  //          l_u = L_B^{-1} U_l (nb x rl), l_v = L_A^{-T} V_l (na x rl);
  //   U^{-1} has the off-diagonal block -u_u u_v^T, rows of A x columns of B,
  //          u_u = U_A^{-1} U_c (na x ru), u_v = U_B^{-T} V_c (nb x ru);
  // both exact (no truncation).  Offsets are within the inverse area, ld =
  // rank; nodes in post-order (children before parents).
  struct inv_node {
    const h_block* node;
    int64_t row0;  // first row of A (B follows)
    int na, nb, rl, ru;
    int64_t l_u, l_v, u_u, u_v;
  };
  bool hodlrShaped() const;
  // Empty when the tree is not HODLR-shaped.
  void inverseNodes(std::vector<inv_node>& out) const;
  // Ops forming the inverses from the factored leaves (trsml / trsmu on
  // identity blocks).  The inverse area must hold identity blocks before the
  // ops run.
  h_ops_tree* generateOps_Invert(const h_index* self, int64_t inv_base, int slot = -1) const;
  // generateOps_Invert plus the low-rank factors of inverseNodes(): products
  // of the children's inverses (already formed, post-order) with the
  // off-diagonal factors, as gemm / gemm_plus ops.  The node factor areas must
  // be zero and the leaf areas identity before the ops run; `tmp` hands out
  // the r x r intermediates (zero-filled temp slot).
  h_ops_tree* generateOps_InvertH(const h_index* self, int64_t inv_base, dev_temp* tmp, int slot = -1) const;

  // Leaves in depth-first order.
  void leaves(std::vector<const h_block*>& out) const;
  // Number of stored elements (dense entries + U and V entries).
  int64_t storedEntries() const;
};
