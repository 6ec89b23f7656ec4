/*
 * thcgw_eri.h - C ABI of libthcgw_eri.so: the THC ERI file (thc.eri.h5) on the host.
 *
 * The reference's THC operator (thc_reader_t) writes the ISDF result to an HDF5 file
 * after a build and reads it back instead of rebuilding: thc_reader_t::read()
 * (src/methods/ERI/thc_reader_t.hpp:507-574) and the write at :236-256 with the
 * builder's datasets (src/methods/ERI/thc.cpp:361-377).  These functions read that file
 * into caller-owned host buffers in the layouts of thcgw.h (X_skPa, V_qPQ), ready to be
 * uploaded for the GPU path, and write the same layout (for converters and tests).
 *
 * Complex arrays are interleaved complex<double>.  V is coulomb_matrix as stored, i.e.
 * thc_reader_t::Z(iq) including the N_k factor.  Each function returns THCGW_ERI_OK or
 * THCGW_ERI_ERR with a message in err (errlen bytes, may be NULL).  Needs no GPU.
 */
#ifndef THCGW_ERI_H
#define THCGW_ERI_H

#ifdef __cplusplus
extern "C" {
#endif

#define THCGW_ERI_OK 0
#define THCGW_ERI_ERR 1

typedef struct thcgw_eri_info {
  int Np;                /* THC points (thc_reader_t::Np) */
  int nbnd;              /* number_of_bands (-1 if absent) */
  int x_first, x_last;   /* X_orbital_range: orbitals [x_first, x_last) */
  int y_first, y_last;   /* Y_orbital_range */
  int ns_npol;           /* first dimension of collocation_matrix */
  int nkpts;             /* k-points of collocation_matrix */
  int nkpts_ibz;         /* nkpts_ibz (nkpts if absent) */
  int nqpts_ibz;         /* first dimension of coulomb_matrix */
  int has_Y;             /* Y_collocation_matrix present */
  int nkpts_file;        /* rows of the kpts dataset (0 if absent; nkpts when present) */
  int nqpts;             /* rows of the qpts dataset (0 if absent) */
} thcgw_eri_info;

/* Sizes and consistency checks of thc_reader_t::read (ranges within nbnd,
 * interpolating_points of length Np, collocation and Coulomb shapes), plus the shapes of
 * Y_collocation_matrix (ns*npol, nkpts, Np, y_last - y_first), kpts (nkpts, 3) and
 * qpts (nqpts, 3). */
int thcgw_eri_read_info(const char* path, thcgw_eri_info* info, char* err, int errlen);
/* The read functions take the capacity of each output buffer in doubles and refuse a
 * dataset of any other size, so a buffer sized from a different file is never overrun. */
/* collocation_matrix (which = 0) or Y_collocation_matrix (which = 1):
 * (ns*npol, nkpts, Np, x_last - x_first) resp. (..., y_last - y_first) complex. */
int thcgw_eri_read_X(const char* path, int which, double* X, long capacity, char* err, int errlen);
/* coulomb_matrix: (nqpts_ibz, Np, Np) complex. */
int thcgw_eri_read_V(const char* path, double* V, long capacity, char* err, int errlen);
/* kpts (nkpts, 3) and qpts (nqpts, 3); either may be NULL. */
int thcgw_eri_read_kpts(const char* path, double* kpts, long kpts_capacity, double* qpts, long qpts_capacity,
                        char* err, int errlen);
/* Write the file; Y is used only when the ranges differ; kpts ((nkpts, 3)), qpts
 * ((info->nqpts, 3), nkpts rows when info->nqpts <= 0), interp_points (Np longs), chi_head
 * and chi_bar_head ((nqpts_ibz, Np) complex) may be NULL. */
int thcgw_eri_write(const char* path, const thcgw_eri_info* info, const double* X, const double* Y,
                    const double* V, const double* kpts, const double* qpts, const long* interp_points,
                    const double* chi_head, const double* chi_bar_head, char* err, int errlen);

#ifdef __cplusplus
}
#endif

#endif /* THCGW_ERI_H */
