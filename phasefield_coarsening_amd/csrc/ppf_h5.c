/*
 * ppf_h5.c -- the HDF5 file contract of the stepping program.
 *
 * Input (phasefield.c:996-1049, README.md:26-34): float64 datasets /etaS,
 * /etaL and /c, C order [Nz][Ny][Nx] with x fastest.  The reference reads the
 * [0,N) corner of each dataset as H5T_IEEE_F64LE into compile-time sized
 * arrays; here the sizes come from /etaS itself and the other datasets must
 * be at least that large (the same corner read).
 *
 * The reference's reader never loads /etaL: it opens the dataset, then reads
 * through the already closed /etaS id (phasefield.c:1026 closes it,
 * phasefield.c:1032 reads with it), so etaL keeps its init value 1.0
 * everywhere (phasefield.c:160).  PPFIO_ETAL_COMPAT reproduces that; without
 * it /etaL is read as README.md:30,33 documents.
 *
 * Output (phasefield.c:1052-1109): a new file (H5F_ACC_TRUNC) with contiguous
 * float64 datasets etaS, etaL, c of dims [Nz][Ny][Nx].
 *
 * All calls return 0 or a negative PPFIO_E* code; ppfio_last_error() gives
 * the message.  Nothing is printed: HDF5's automatic error stack printing is
 * switched off.
 */
#include <hdf5.h>
#include <stdio.h>
#include <string.h>

#include "ppf_io.h"

static __thread char g_err[512];

static int fail(int code, const char *fmt, const char *a, const char *b)
{
    snprintf(g_err, sizeof g_err, fmt, a ? a : "", b ? b : "");
    return code;
}

const char *ppfio_last_error(void) { return g_err; }

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

/* dims of a 3-D dataset; returns 0 or PPFIO_E* */
static int ds_dims(hid_t file, const char *name, hsize_t d[3])
{
    if (H5Lexists(file, name, H5P_DEFAULT) <= 0)
        return fail(PPFIO_EMISSING, "dataset %s missing%s", name, NULL);
    hid_t ds = H5Dopen2(file, name, H5P_DEFAULT);
    if (ds < 0)
        return fail(PPFIO_EMISSING, "cannot open dataset %s%s", name, NULL);
    hid_t sp = H5Dget_space(ds);
    int nd = H5Sget_simple_extent_ndims(sp);
    int rc = 0;
    if (nd != 3)
        rc = fail(PPFIO_EDIMS, "dataset %s is not 3-D%s", name, NULL);
    else
        H5Sget_simple_extent_dims(sp, d, NULL);
    H5Sclose(sp);
    H5Dclose(ds);
    return rc;
}

int ppfio_dims(const char *path, int dims[3])
{
    quiet();
    g_err[0] = 0;
    if (!path || !dims)
        return fail(PPFIO_EARG, "null argument%s%s", NULL, NULL);
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0)
        return fail(PPFIO_EOPEN, "cannot open %s for reading%s", path, NULL);
    hsize_t d[3];
    int rc = ds_dims(f, "/etaS", d);
    H5Fclose(f);
    if (rc)
        return rc;
    if (d[0] < 1 || d[1] < 1 || d[2] < 1 || d[0] > (1u << 30) || d[1] > (1u << 30) ||
        d[2] > (1u << 30))
        return fail(PPFIO_EDIMS, "dataset /etaS has unusable dims%s%s", NULL, NULL);
    dims[0] = (int)d[0];
    dims[1] = (int)d[1];
    dims[2] = (int)d[2];
    return 0;
}

/* read the [0,dims) corner of one dataset into a dense buffer */
static int read_corner(hid_t f, const char *name, const int dims[3], double *out)
{
    hsize_t d[3];
    int rc = ds_dims(f, name, d);
    if (rc)
        return rc;
    for (int a = 0; a < 3; a++)
        if (d[a] < (hsize_t)dims[a]) {
            char buf[64];
            snprintf(buf, sizeof buf, "%d/%d/%d", dims[0], dims[1], dims[2]);
            return fail(PPFIO_EDIMS, "dataset %s is smaller than %s", name, buf);
        }
    hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
    hid_t fsp = H5Dget_space(ds);
    hsize_t off[3] = {0, 0, 0}, cnt[3] = {(hsize_t)dims[0], (hsize_t)dims[1], (hsize_t)dims[2]};
    H5Sselect_hyperslab(fsp, H5S_SELECT_SET, off, NULL, cnt, NULL);
    hid_t msp = H5Screate_simple(3, cnt, NULL);
    herr_t st = H5Dread(ds, H5T_IEEE_F64LE, msp, fsp, H5P_DEFAULT, out);
    H5Sclose(msp);
    H5Sclose(fsp);
    H5Dclose(ds);
    if (st < 0)
        return fail(PPFIO_EREAD, "reading dataset %s failed%s", name, NULL);
    return 0;
}

int ppfio_read(const char *path, const int dims[3], double *s, double *l, double *c, int flags)
{
    quiet();
    g_err[0] = 0;
    if (!path || !dims || !s || !l || !c)
        return fail(PPFIO_EARG, "null argument%s%s", NULL, NULL);
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0)
        return fail(PPFIO_EOPEN, "cannot open %s for reading%s", path, NULL);
    int rc = read_corner(f, "/etaS", dims, s);
    if (!rc) {
        if (flags & PPFIO_ETAL_COMPAT) {
            size_t n = (size_t)dims[0] * dims[1] * dims[2];
            for (size_t q = 0; q < n; q++)
                l[q] = 1.0;
        } else {
            rc = read_corner(f, "/etaL", dims, l);
        }
    }
    if (!rc)
        rc = read_corner(f, "/c", dims, c);
    H5Fclose(f);
    return rc;
}

int ppfio_write(const char *path, const int dims[3], const double *s, const double *l,
                const double *c)
{
    return ppfio_write_chunked(path, dims, s, l, c, NULL);
}

/* As ppfio_write, with chunked datasets when chunk is given: the layout of
 * the reference's parallel writer, which chunks each dataset by the local
 * block of its MPI decomposition (phasefield.c:576-582).  The values and
 * dataset shapes are those of the contiguous file; only the storage layout
 * differs. */
int ppfio_write_chunked(const char *path, const int dims[3], const double *s, const double *l,
                        const double *c, const int chunk[3])
{
    quiet();
    g_err[0] = 0;
    if (!path || !dims || !s || !l || !c)
        return fail(PPFIO_EARG, "null argument%s%s", NULL, NULL);
    if (chunk && (chunk[0] == 0 && chunk[1] == 0 && chunk[2] == 0))
        chunk = NULL;
    if (chunk)
        for (int i = 0; i < 3; i++)
            if (chunk[i] < 1 || chunk[i] > dims[i])
                return fail(PPFIO_EARG, "chunk dims must lie in 1..dims%s%s", NULL, NULL);
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0)
        return fail(PPFIO_EOPEN, "cannot create %s%s", path, NULL);
    hsize_t d[3] = {(hsize_t)dims[0], (hsize_t)dims[1], (hsize_t)dims[2]};
    hid_t sp = H5Screate_simple(3, d, NULL);
    hid_t dcpl = H5P_DEFAULT;
    if (chunk) {
        hsize_t cd[3] = {(hsize_t)chunk[0], (hsize_t)chunk[1], (hsize_t)chunk[2]};
        dcpl = H5Pcreate(H5P_DATASET_CREATE);
        H5Pset_chunk(dcpl, 3, cd);
    }
    const char *names[3] = {"etaS", "etaL", "c"};
    const double *bufs[3] = {s, l, c};
    int rc = 0;
    for (int a = 0; a < 3 && !rc; a++) {
        hid_t ds = H5Dcreate2(f, names[a], H5T_IEEE_F64LE, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT);
        if (ds < 0) {
            rc = fail(PPFIO_EWRITE, "cannot create dataset %s in %s", names[a], path);
            break;
        }
        if (H5Dwrite(ds, H5T_IEEE_F64LE, H5S_ALL, H5S_ALL, H5P_DEFAULT, bufs[a]) < 0)
            rc = fail(PPFIO_EWRITE, "writing dataset %s in %s failed", names[a], path);
        H5Dclose(ds);
    }
    if (dcpl != H5P_DEFAULT)
        H5Pclose(dcpl);
    H5Sclose(sp);
    if (H5Fclose(f) < 0 && !rc)
        rc = fail(PPFIO_EWRITE, "closing %s failed%s", path, NULL);
    return rc;
}

/* Storage layout of a 3-D dataset: chunk dims, or zeros when contiguous. */
int ppfio_layout(const char *path, const char *dataset, int chunk[3])
{
    quiet();
    g_err[0] = 0;
    if (!path || !dataset || !chunk)
        return fail(PPFIO_EARG, "null argument%s%s", NULL, NULL);
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0)
        return fail(PPFIO_EOPEN, "cannot open %s for reading%s", path, NULL);
    int rc = 0;
    hid_t ds = H5Dopen2(f, dataset, H5P_DEFAULT);
    if (ds < 0) {
        rc = fail(PPFIO_EMISSING, "cannot open dataset %s%s", dataset, NULL);
    } else {
        hid_t dcpl = H5Dget_create_plist(ds);
        chunk[0] = chunk[1] = chunk[2] = 0;
        if (H5Pget_layout(dcpl) == H5D_CHUNKED) {
            hsize_t cd[3] = {0, 0, 0};
            if (H5Pget_chunk(dcpl, 3, cd) != 3)
                rc = fail(PPFIO_EDIMS, "dataset %s chunk is not 3-D%s", dataset, NULL);
            for (int i = 0; i < 3; i++)
                chunk[i] = (int)cd[i];
        }
        H5Pclose(dcpl);
        H5Dclose(ds);
    }
    H5Fclose(f);
    return rc;
}

/* Job stem, phasefield.c:117-120: the input name with its last three
 * characters replaced by one '_' ("run.h5" -> "run_").  The reference copies
 * into char[50] unchecked; here the caller gives the capacity. */
int ppfio_job_stem(const char *input, char *stem, int cap)
{
    g_err[0] = 0;
    if (!input || !stem || cap < 1)
        return fail(PPFIO_EARG, "null argument%s%s", NULL, NULL);
    size_t n = strlen(input);
    if (n < 3)
        return fail(PPFIO_EARG, "input name %s is shorter than 3 characters%s", input, NULL);
    if (n - 1 >= (size_t)cap)
        return fail(PPFIO_EARG, "input name %s too long%s", input, NULL);
    memcpy(stem, input, n - 3);
    stem[n - 3] = '_';
    stem[n - 2] = '\0';
    return 0;
}

/* Output name, phasefield.c:1193-1206 (itoa of the save counter, then
 * ".h5"): stem "run_" and k = 3 give "run_3.h5". */
int ppfio_output_name(const char *stem, int k, char *out, int cap)
{
    g_err[0] = 0;
    if (!stem || !out || cap < 1)
        return fail(PPFIO_EARG, "null argument%s%s", NULL, NULL);
    int n = snprintf(out, (size_t)cap, "%s%d.h5", stem, k);
    if (n < 0 || n >= cap)
        return fail(PPFIO_EARG, "output name for stem %s too long%s", stem, NULL);
    return 0;
}
