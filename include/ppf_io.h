/*
 * ppf_io.h -- HDF5 file contract of the stepping program (libppfio.so).
 *
 * Replaces the reference's serial reader/writer and output naming:
 *   ppfio_read        read_hdf5            phasefield.c:996-1049
 *   ppfio_write       write_hdf5           phasefield.c:1052-1109
 *   ppfio_write_chunked  write_hdf5_parallel's chunked layout  phasefield.c:548-624
 *   ppfio_job_stem    argv[1] -> job name  phasefield.c:117-121
 *   ppfio_output_name write_result_file    phasefield.c:1193-1206 (itoa 285-299)
 * ppfio_dims has no counterpart: the reference fixes sizes at compile time
 * (phasefield.c:47-56) and this library takes them from /etaS.
 *
 * Arrays are dense C order [dims[0]][dims[1]][dims[2]] = [Nz][Ny][Nx], x
 * fastest, float64, caller-owned.  Every call returns 0 or a negative
 * PPFIO_E* code and never prints.
 */
#ifndef PPF_IO_H
#define PPF_IO_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PPFIO_OK = 0,
    PPFIO_EARG = -1,
    PPFIO_EOPEN = -2,
    PPFIO_EMISSING = -3,
    PPFIO_EDIMS = -4,
    PPFIO_EREAD = -5,
    PPFIO_EWRITE = -6
};

/* ppfio_read flag: ignore /etaL and set etaL = 1.0, as the reference's reader
 * does (phasefield.c:1032 reads through the closed /etaS id). */
#define PPFIO_ETAL_COMPAT 1

int ppfio_dims(const char *path, int dims[3]);
int ppfio_read(const char *path, const int dims[3], double *s, double *l, double *c, int flags);
int ppfio_write(const char *path, const int dims[3], const double *s, const double *l,
                const double *c);
/* ppfio_write with datasets chunked by chunk[3] (z, y, x), the on-disk
 * layout of the reference's parallel writer, write_hdf5_parallel, whose
 * chunk is the local block of its MPI decomposition (phasefield.c:548-624).
 * chunk NULL or all zero: contiguous, as ppfio_write. */
int ppfio_write_chunked(const char *path, const int dims[3], const double *s, const double *l,
                        const double *c, const int chunk[3]);
/* Chunk dims of a dataset, all zero when it is stored contiguously. */
int ppfio_layout(const char *path, const char *dataset, int chunk[3]);
int ppfio_job_stem(const char *input, char *stem, int cap);
int ppfio_output_name(const char *stem, int k, char *out, int cap);
const char *ppfio_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
