/*
 * mdsctk_graph.h - the stages after knn_rms: the symmetric sparse kNN
 * graph (make_sysparse) and its leading spectral decomposition
 * (decomp_sparse); the general sparse matrix of out-of-sample rows
 * (make_gesparse) and the Nystrom extension of the eigenvectors to them
 * (decomp_sparse_nystrom).
 *
 * .ssm layout (make_sysparse.cpp:285-316, read back by CSC_matrix at
 * mdsctk.cpp:44-59 and decomp_sparse.cpp:137-145), host byte order,
 * headerless:
 *     int32 n; int32 pcol[n+1]; int32 irow[nnz]; float64 vals[nnz]
 * with nnz = pcol[n].  Column c holds the edges (c, r), r > c, of the
 * symmetrised kNN relation, rows ascending: the strict lower triangle of a
 * symmetric matrix in CSC form.  Values are the kNN distances (Angstrom).
 *
 * Return codes as in mdsctk_knn.h (0 ok, negative MDSCTK_ERR_*); the text
 * of the last failure is mdsctk_io_last_error().
 */
#ifndef MDSCTK_GRAPH_H
#define MDSCTK_GRAPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Symmetric sparse matrix from kNN rows, make_sysparse semantics
 * (make_sysparse.cpp:226-279, compare_edge / split_edges at
 * mdsctk.cpp:567-604): for every row i and its first k of maxk neighbours
 * j (dist/idx rows of width maxk), the edge (min(i,j), max(i,j)) gets that
 * distance; a later row overwrites an earlier value of the same edge;
 * i == j is skipped.  n = n_rows.
 * Call with pcol/irow/vals NULL to get *nnz; then with buffers of n_rows+1
 * and *nnz entries. */
int mdsctk_ssm_build(const double* dist, const int32_t* idx, int64_t n_rows, int32_t maxk,
                     int32_t k, int32_t* pcol, int32_t* irow, double* vals, int64_t* nnz);

/* Write / read an .ssm file.  mdsctk_ssm_read with NULL buffers returns n
 * and nnz only. */
int mdsctk_ssm_write(const char* path, int32_t n, const int32_t* pcol, const int32_t* irow,
                     const double* vals);
int mdsctk_ssm_read(const char* path, int32_t* n, int64_t* nnz, int32_t* pcol, int32_t* irow,
                    double* vals);

/* decomp_sparse (decomp_sparse.cpp:147-277): affinities
 * a = exp(-d^2 / (2 sigma^2)) on the stored edges, degrees d_i = sum of
 * row i of the symmetric matrix (no diagonal), normalisation
 * D^-1/2 A D^-1/2, then the nev largest algebraic eigenvalues with their
 * eigenvectors.  Thick-restart Lanczos with full reorthogonalisation and a
 * subspace of ncv = min(10 nev + 1, n) vectors (the reference's ARPACK
 * setting), converged to machine precision.
 *   evals     [nev]      descending (decomp_sparse writes them largest first)
 *   evecs     [nev][n]   unit eigenvectors (sign arbitrary), may be NULL
 *   residuals [nev]      ||A z - l z|| / |l|, may be NULL
 * Returns the number of converged eigenpairs (== nev on success). */
int mdsctk_spectral_decomp(int32_t n, const int32_t* pcol, const int32_t* irow,
                           const double* dist, double sigma, int32_t nev, double* evals,
                           double* evecs, double* residuals);

/* General sparse matrix of out-of-sample kNN rows, make_gesparse semantics
 * (make_gesparse.cpp:226-333, split_edges at mdsctk.cpp:580-603).  The .gsm
 * file has the .ssm layout (mdsctk_ssm_write / mdsctk_ssm_read handle it)
 * with one column per query row: column i holds the entries (i, j) for the
 * first k of maxk neighbours j of row i, rows ascending, values the kNN
 * distances; j indexes the reference set and may exceed n_rows; i == j is
 * kept.  A row's own entry replaces an earlier one of the same edge.
 * With symmetric != 0 (make_gesparse -s) the mirrored entry (j, i) is added
 * when that edge has no entry yet; mirrored entries with j >= n_rows have no
 * column and are dropped, as in the reference's conversion loop.
 * Call with pcol/irow/vals NULL to get *nnz; then with buffers of n_rows+1
 * and *nnz entries. */
int mdsctk_gsm_build(const double* dist, const int32_t* idx, int64_t n_rows, int32_t maxk,
                     int32_t k, int symmetric, int32_t* pcol, int32_t* irow, double* vals,
                     int64_t* nnz);

/* decomp_sparse_nystrom (decomp_sparse_nystrom.cpp:147-304): the leading
 * eigenpairs of the in-sample .ssm matrix (n x n) as mdsctk_spectral_decomp,
 * plus their Nystrom extension to the m out-of-sample columns of a .gsm
 * matrix whose row indices lie in [0, n):
 *   b_ij = exp(-d_ij^2 / 2 sigma^2) * d_i^-1/2 * (sum_i' b_i'j)^-1/2
 *   ext[x][j] = sum_i b_ij z_x[i] / l_x
 * with d_i the in-sample degrees.  evecs [nev][n] may be NULL; ext is
 * [nev][m].  Returns the number of converged eigenpairs (== nev on
 * success). */
int mdsctk_nystrom_decomp(int32_t n, const int32_t* pcol_a, const int32_t* irow_a,
                          const double* dist_a, int32_t m, const int32_t* pcol_b,
                          const int32_t* irow_b, const double* dist_b, double sigma, int32_t nev,
                          double* evals, double* evecs, double* ext, double* residuals);

#ifdef __cplusplus
}
#endif

#endif /* MDSCTK_GRAPH_H */
