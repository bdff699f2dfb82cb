// Host ordering for the spaQR tree: geometric nested dissection on the graph
// of A^T A and the (self, left, right) interface labels of every column.
#pragma once
#include <vector>

#include "common.hpp"

namespace spq {

// Compressed sparse column matrix, 0-based, sorted row indices per column.
struct CscMat {
  int m = 0, n = 0;
  std::vector<int> p, i;
  std::vector<double> x;
  int nnz() const { return (int)i.size(); }
};
CscMat csc_transpose(const CscMat& A);
CscMat csc_permute(const CscMat& A, const std::vector<int>* rperm, const std::vector<int>* cperm);

// Labels per column (size n, or m for rectangular A where the no-matching
// rule copies row labels onto columns).  Follows partition.cpp:241-348 and
// 506-578 with the no-HSL row/column matching (partition.cpp:149-156).
std::vector<CID> geometric_labels(const CscMat& A, int nlevels, const std::vector<double>& X, int dim);

// tree.cpp:46-55: stable sorts of the column order by the labels merged level
// by level.
void order_columns(int nlevels, const std::vector<CID>& labels, std::vector<int>& cperm);

// tree.cpp:100-163 (no matching): rows grouped behind the column of their
// heaviest cluster group; returns the row order and the rows per column.
void order_rows_ls(const CscMat& Ap, const std::vector<CID>& cpermed, std::vector<int>& rperm,
                   std::vector<int>& rows_per_col);

std::vector<double> tensor_coordinates(int n, int dim);  // util.cpp:852-881

// Non-geometric path (tree.cpp:608-617).  row_parts: the library's own
// partition of the rows of A into 2^(nlevels-1) parts (recursive bisection
// of the column-net hypergraph), in place of partition.cpp:422-448 (PaToH);
// parts 2g and 2g+1 are the halves of part g one level up.
std::vector<int> row_parts(const CscMat& A, int nlevels);
// Column labels from a row partition: partition.cpp:450-499 (separators,
// leaves) and 582-623 (getInterfacesUsingA_HUND).
std::vector<CID> hypergraph_labels(const CscMat& A, int nlevels, const std::vector<int>& parts);

}  // namespace spq
