/*
 * C interface of the MI355X H-matrix engine.
 *
 * This header is ABI-compatible with hmat-oss's public header
 * (include/hmat/hmat.h of the reference): the enums carry the same values,
 * the structs the same members in the same order, and hmat_interface_t the
 * same function-pointer slots in the same order (hmat.h:717-1103).  An
 * application written against hmat-oss selects this engine by calling
 * hmat_init_mi355x_interface() where it called hmat_init_default_interface();
 * the rest of its code is unchanged.
 *
 * Entry points whose slot exists but that the engine does not provide are set
 * to stubs that print a message on stderr and return 1 (or NULL).
 */
#ifndef HMAT_MI355X_HMAT_H
#define HMAT_MI355X_HMAT_H

#include <stdlib.h>
#include <stdbool.h>

#define HMAT_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  HMAT_SIMPLE_PRECISION = 0,
  HMAT_DOUBLE_PRECISION = 1,
  HMAT_SIMPLE_COMPLEX = 2,
  HMAT_DOUBLE_COMPLEX = 3
} hmat_value_t;

typedef enum {
  hmat_compress_svd,
  hmat_compress_aca_full,
  hmat_compress_aca_partial,
  hmat_compress_aca_plus,
  hmat_compress_aca_random,
  hmat_compress_rrqr
} hmat_compress_t;

typedef enum {
  DEFAULT_COMPRESSOR = 0,
  ZFP_COMPRESSOR = 1,
  SZ_COMPRESSOR = 2,
  SZ3_COMPRESSOR = 3
} hmat_FPcompress_t;

typedef enum { hmat_block_full, hmat_block_null, hmat_block_sparse } hmat_block_t;

typedef enum {
  hmat_factorization_none = -1,
  hmat_factorization_lu,
  hmat_factorization_ldlt,
  hmat_factorization_llt,
  hmat_factorization_hodlr,
  hmat_factorization_hodlrsym
} hmat_factorization_t;

typedef struct hmat_block_info_struct {
  hmat_block_t block_type;
  void *user_data;
  void (*release_user_data)(void *user_data);
  char (*is_guaranteed_null_row)(const struct hmat_block_info_struct *block_info,
                                 int block_row_offset, int stratum);
  char (*is_guaranteed_null_col)(const struct hmat_block_info_struct *block_info,
                                 int block_col_offset, int stratum);
  size_t needed_memory;
  int number_of_strata;
} hmat_block_info_t;

typedef struct hmat_fp_settings_struct {
  hmat_FPcompress_t compressor;
  int nb_blocs;
  float epsilonFP;
  bool compressFull;
  bool compressRk;
} hmat_fp_settings_t;

typedef void (*hmat_prepare_func_t)(int row_start, int row_count, int col_start, int col_count,
                                    int *row_hmat2client, int *row_client2hmat,
                                    int *col_hmat2client, int *col_client2hmat, void *context,
                                    hmat_block_info_t *block_info);

typedef void (*hmat_compute_func_t)(void *v_data, int block_row_start, int block_row_count,
                                    int block_col_start, int block_col_count, void *block);

/* One matrix entry in user numbering; writes a double (real types) or a
 * double _Complex (complex types) to *result. */
typedef void (*hmat_interaction_func_t)(void *user_context, int row, int col, void *result);

typedef struct hmat_clustering_algorithm hmat_clustering_algorithm_t;
typedef struct hmat_cluster_tree_struct hmat_cluster_tree_t;

HMAT_API hmat_clustering_algorithm_t *hmat_create_clustering_median(void);
HMAT_API hmat_clustering_algorithm_t *hmat_create_clustering_max_dof(
    const hmat_clustering_algorithm_t *algo, int max_dof);
HMAT_API void hmat_delete_clustering(hmat_clustering_algorithm_t *algo);
HMAT_API void hmat_set_clustering_divider(hmat_clustering_algorithm_t *algo, int divider);

HMAT_API hmat_cluster_tree_t *hmat_create_cluster_tree(double *coord, int dimension, int size,
                                                       hmat_clustering_algorithm_t *algo);

typedef struct hmat_cluster_tree_builder hmat_cluster_tree_builder_t;
HMAT_API hmat_cluster_tree_builder_t *hmat_create_cluster_tree_builder(
    const hmat_clustering_algorithm_t *algo);
HMAT_API void hmat_delete_cluster_tree_builder(hmat_cluster_tree_builder_t *ctb);

struct hmat_cluster_tree_create_context_t {
  unsigned dimension;
  unsigned number_of_points;
  double *coordinates;
  unsigned number_of_dof;
  unsigned *span_offsets;
  unsigned *spans;
  int *group_index;
  const hmat_cluster_tree_builder_t *builder;
};
HMAT_API hmat_cluster_tree_t *hmat_create_cluster_tree_generic(
    struct hmat_cluster_tree_create_context_t *);

HMAT_API void hmat_delete_cluster_tree(const hmat_cluster_tree_t *tree);
HMAT_API int hmat_tree_nodes_count(const hmat_cluster_tree_t *tree);

typedef struct {
  int spatial_dimension;
  int dimension;
  size_t nr_tree_nodes;
} hmat_cluster_info_t;
HMAT_API int hmat_cluster_get_info(hmat_cluster_tree_t *tree, hmat_cluster_info_t *info);
HMAT_API const int *hmat_cluster_get_indices(const hmat_cluster_tree_t *tree);

typedef struct {
  double eta;
  double ratio;
  size_t max_width;
  double scale;
} hmat_admissibility_param_t;

HMAT_API void hmat_init_admissibility_param(hmat_admissibility_param_t *);
typedef struct hmat_admissibility_condition hmat_admissibility_t;
HMAT_API hmat_admissibility_t *hmat_create_admissibility(hmat_admissibility_param_t *);
HMAT_API void hmat_update_admissibility(hmat_admissibility_t *, hmat_admissibility_param_t *);
HMAT_API hmat_admissibility_t *hmat_create_admissibility_standard(double eta);
/* Admissibility of a HODLR matrix: every block between two different
 * clusters is low rank (hmat.h:358-362 of hmat-oss). */
HMAT_API hmat_admissibility_t *hmat_create_admissibility_hodlr(void);
HMAT_API void hmat_delete_admissibility(hmat_admissibility_t *cond);

typedef struct hmat_compression_algorithm_struct hmat_compression_algorithm_t;
HMAT_API hmat_compression_algorithm_t *hmat_create_compression_svd(double epsilon);
HMAT_API hmat_compression_algorithm_t *hmat_create_compression_aca_partial(double epsilon);
HMAT_API hmat_compression_algorithm_t *hmat_create_compression_aca_plus(double epsilon);
HMAT_API void hmat_delete_compression(const hmat_compression_algorithm_t *algo);

typedef struct {
  int full_zeros;
  size_t full_size;
  size_t rk_size;
  size_t full_count;
  size_t rk_count;
  size_t compressed_size;
  size_t uncompressed_size;
  int nr_block_clusters;
  int largest_rk_dim_rows;
  int largest_rk_dim_cols;
  int largest_rk_mem_rows;
  int largest_rk_mem_cols;
  int largest_rk_mem_rank;
} hmat_info_t;

typedef struct {
  int wip;
} hmat_profile_t;

typedef struct hmat_matrix_struct hmat_matrix_t;

typedef struct hmat_progress_struct {
  int max;
  int current;
  void (*update)(struct hmat_progress_struct *context);
  void *user_data;
} hmat_progress_t;

HMAT_API hmat_progress_t *hmat_default_progress(void);

typedef struct {
  double rkRatio;
  size_t size_Rk;
  size_t size_Rk_compressed;
  double fullRatio;
  size_t size_Full;
  size_t size_Full_compressed;
  double ratio;
} hmat_FPCompressionRatio_t;

typedef void (*hmat_iostream)(void *buffer, size_t n, void *user_data);

struct hmat_block_compute_context_t {
  void *user_data;
  int row_start, row_count, col_start, col_count;
  int stratum;
  void *block;
};

typedef struct {
  void *assembly;
  hmat_interaction_func_t simple_compute;
  hmat_compute_func_t block_compute;
  void (*advanced_compute)(struct hmat_block_compute_context_t *);
  void *user_context;
  hmat_prepare_func_t prepare;
  const hmat_compression_algorithm_t *compression;
  int lower_symmetric;
  hmat_factorization_t factorization;
  hmat_progress_t *progress;
} hmat_assemble_context_t;

HMAT_API void hmat_assemble_context_init(hmat_assemble_context_t *context);

typedef struct {
  hmat_factorization_t factorization;
  hmat_progress_t *progress;
} hmat_factorization_context_t;

HMAT_API void hmat_factorization_context_init(hmat_factorization_context_t *context);

struct hmat_get_values_context_t {
  hmat_matrix_t *matrix;
  void *values;
  int row_offset, col_offset;
  int row_size, col_size;
  int *row_indices, *col_indices;
  int renumber_rows : 1;
};

struct hmat_solve_context_t {
  void *values;
  int nr_rhs;
  int no_permutation;
  int lower;
  int upper;
  hmat_progress_t *progress;
};

HMAT_API void hmat_solve_context_init(struct hmat_solve_context_t *context);

typedef struct {
  hmat_value_t value_type;
  void *internal;
} hmat_procedure_t;

typedef struct {
  hmat_value_t value_type;
  const void *internal;
} hmat_leaf_procedure_t;

/* Function table; slot order and signatures follow hmat.h:717-1103. */
typedef struct {
  hmat_matrix_t *(*create_empty_hmatrix_admissibility)(const hmat_cluster_tree_t *rows_tree,
                                                       const hmat_cluster_tree_t *cols_tree,
                                                       int lower_symmetric,
                                                       hmat_admissibility_t *cond);
  void (*own_cluster_trees)(hmat_matrix_t *hmatrix, int owns_row, int owns_col);
  void (*set_low_rank_epsilon)(hmat_matrix_t *hmatrix, double epsilon);
  int (*assemble_generic)(hmat_matrix_t *matrix, hmat_assemble_context_t *context);
  hmat_matrix_t *(*copy)(hmat_matrix_t *hmatrix);
  hmat_matrix_t *(*copy_struct)(hmat_matrix_t *hmatrix);
  double (*norm)(hmat_matrix_t *hmatrix);
  int (*logdet)(hmat_matrix_t *hmatrix, double *result);
  int (*inverse)(hmat_matrix_t *hmatrix);
  int (*factorize_generic)(hmat_matrix_t *matrix, hmat_factorization_context_t *context);
  int (*destroy)(hmat_matrix_t *hmatrix);
  hmat_matrix_t *(*get_child)(hmat_matrix_t *hmatrix, int i, int j);
  int (*destroy_child)(hmat_matrix_t *hmatrix);
  int (*solve_generic)(hmat_matrix_t *matrix, const struct hmat_solve_context_t *context);
  int (*solve_mat)(hmat_matrix_t *hmatrix, hmat_matrix_t *hmatrixB);
  int (*solve_systems)(hmat_matrix_t *hmatrix, void *b, int nrhs);
  int (*solve_dense)(hmat_matrix_t *hmatrix, void *b, int nrhs);
  int (*transpose)(hmat_matrix_t *hmatrix);
  int (*scale)(void *alpha, hmat_matrix_t *hmatrix);
  int (*truncate)(hmat_matrix_t *hmatrix);
  int (*vector_reorder)(void *vec_b, const hmat_cluster_tree_t *rols_ct, int rows,
                        const hmat_cluster_tree_t *cols_ct, int cols);
  int (*vector_restore)(void *vec_b, const hmat_cluster_tree_t *rols_ct, int rows,
                        const hmat_cluster_tree_t *cols_ct, int cols);
  int (*gemm)(char trans_a, char trans_b, const void *alpha, hmat_matrix_t *hmatrix,
              hmat_matrix_t *hmatrix_b, const void *beta, hmat_matrix_t *hmatrix_c);
  int (*axpy)(void *a, hmat_matrix_t *x, hmat_matrix_t *hmatrix_y);
  int (*trsm)(char side, char uplo, char transa, char diag, int m, int n, void *alpha,
              hmat_matrix_t *A, int is_b_hmat, void *B);
  int (*gemv)(char trans_a, void *alpha, hmat_matrix_t *hmatrix, void *vec_b, void *beta,
              void *vec_c, int nrhs);
  int (*gemm_scalar)(char trans_a, void *alpha, hmat_matrix_t *hmatrix, void *vec_b, void *beta,
                     void *vec_c, int nrhs);
  int (*full_gemm)(char trans_a, char trans_b, int mc, int nc, void *c, void *alpha, void *a,
                   hmat_matrix_t *hmat_b, void *beta);
  int (*gemm_dense)(char trans_b, char trans_x, char side, const void *alpha,
                    hmat_matrix_t *holder, void *vec_x, const void *beta, void *vec_y,
                    int nrhs);
  int (*add_identity)(hmat_matrix_t *hmatrix, void *alpha);
  int (*init)(void);
  int (*finalize)(void);
  int (*get_info)(hmat_matrix_t *hmatrix, hmat_info_t *info);
  int (*get_profile)(hmat_matrix_t *hmatrix, hmat_profile_t *profile);
  int (*get_ratio)(hmat_matrix_t *hmatrix, hmat_FPCompressionRatio_t *ratio);
  int (*FPcompress)(hmat_matrix_t *hmatrix);
  int (*FPdecompress)(hmat_matrix_t *hmatrix);
  hmat_fp_settings_t (*GetFPCompressionSettings)(hmat_matrix_t *hmatrix);
  int (*SetFPCompressionSettings)(hmat_matrix_t *hmatrix, hmat_fp_settings_t settings);
  int (*SetFPCompressionSettingsParams)(hmat_matrix_t *hmatrix, float epsilonFP, int nb_blocs,
                                        hmat_FPcompress_t compressor, bool compressFull,
                                        bool compressRk);
  int (*dump_info)(hmat_matrix_t *hmatrix, char *prefix);
  int (*get_cluster_trees)(hmat_matrix_t *hmatrix, const hmat_cluster_tree_t **rows,
                           const hmat_cluster_tree_t **cols);
  int (*set_cluster_trees)(hmat_matrix_t *hmatrix, const hmat_cluster_tree_t *rows,
                           const hmat_cluster_tree_t *cols);
  int (*extract_diagonal)(hmat_matrix_t *holder, void *diag, int size);
  int (*solve_lower_triangular)(hmat_matrix_t *hmatrix, int transpose, void *b, int nrhs);
  int (*solve_lower_triangular_dense)(hmat_matrix_t *hmatrix, int transpose, void *b, int nrhs);
  int (*get_block)(struct hmat_get_values_context_t *ctx);
  int (*get_values)(struct hmat_get_values_context_t *ctx);
  int (*walk)(hmat_matrix_t *hmatrix, hmat_procedure_t *proc);
  int (*apply_on_leaf)(hmat_matrix_t *hmatrix, const hmat_leaf_procedure_t *proc);
  hmat_value_t value_type;
  void *internal;
  hmat_matrix_t *(*read_struct)(hmat_iostream readfunc, void *user_data);
  void (*write_struct)(hmat_matrix_t *matrix, hmat_iostream writefunc, void *user_data);
  void (*read_data)(hmat_matrix_t *matrix, hmat_iostream readfunc, void *user_data);
  void (*write_data)(hmat_matrix_t *matrix, hmat_iostream writefunc, void *user_data);
  void (*set_progressbar)(hmat_matrix_t *matrix, hmat_progress_t *progress);
  int (*extract_diagonal_block)(hmat_matrix_t *holder, int components, void *diag);
} hmat_interface_t;

typedef struct {
  int compressionMinLeafSize;
  double coarseningEpsilon;
  int maxLeafSize;
  int coarsening;
  int validateNullRowCol;
  int validateCompression;
  int validateRecompression;
  int dumpTrace;
  int validationReRun;
  int validationDump;
  double validationErrorThreshold;
} hmat_settings_t;

HMAT_API void hmat_get_parameters(hmat_settings_t *settings);
HMAT_API int hmat_set_parameters(hmat_settings_t *);
HMAT_API const char *hmat_get_version(void);

#ifdef __cplusplus
}
#endif

#endif
