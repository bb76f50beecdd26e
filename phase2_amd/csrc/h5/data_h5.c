/*
 * simul.h5 data layer of ph2run over the HDF5 C library
 * (include/ph2run/data.h; behaviour of ph2run/data.c of the reference,
 * file layout of its doc/simul-h5-specs.md).  One process does all I/O:
 * the reference's rank-0 reads followed by MPI broadcasts reduce to the
 * reads.  Built as libphase2_h5.so, linked to libphase2_hip.so for the
 * carrier types' allocators (circ_hamil_init, circ_muldet_init, paulis_set)
 * and the log.
 */
#define LOG_SUBSYS "data"

#include <complex.h>
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hdf5.h"

#include "log.h"
#include "ph2run/data.h"
#include "phase2_h5.h"
#include "phase2/circ.h"
#include "phase2/paulis.h"

#define G_STPREP "state_prep"
#define G_MULTIDET "state_prep/multidet"
#define G_COEFFMAT "state_prep/coeff_matrix"
#define G_HAMIL "pauli_hamil"
#define D_VALUES "values"

/* -- small helpers --------------------------------------------------------- */

static hid_t grp_open(data_id fid, const char *name)
{
	if (fid == DATA_INVALID_FID || fid == 0)
		return H5I_INVALID_HID;
	const hid_t g = H5Gopen2((hid_t)fid, name, H5P_DEFAULT);
	if (g == H5I_INVALID_HID)
		log_error("%s: no such group", name);
	return g;
}

/* Extents of dataset `name` under loc: rank must equal `rank`. */
static int dset_dims(hid_t loc, const char *name, int rank, hsize_t *dims)
{
	const hid_t d = H5Dopen2(loc, name, H5P_DEFAULT);
	if (d == H5I_INVALID_HID) {
		log_error("dataset %s missing", name);
		return -1;
	}
	const hid_t sp = H5Dget_space(d);
	const int r = sp == H5I_INVALID_HID ? -1 :
		H5Sget_simple_extent_ndims(sp);
	int rc = -1;
	if (r != rank)
		log_error("dataset %s: rank %d, expected %d", name, r, rank);
	else if (H5Sget_simple_extent_dims(sp, dims, NULL) == rank)
		rc = 0;
	if (sp != H5I_INVALID_HID)
		H5Sclose(sp);
	H5Dclose(d);
	return rc;
}

static int dset_read(hid_t loc, const char *name, hid_t mem_type, void *buf)
{
	const hid_t d = H5Dopen2(loc, name, H5P_DEFAULT);
	if (d == H5I_INVALID_HID) {
		log_error("dataset %s missing", name);
		return -1;
	}
	const herr_t e = H5Dread(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT,
		buf);
	H5Dclose(d);
	if (e < 0)
		log_error("dataset %s: read failed", name);
	return e < 0 ? -1 : 0;
}

static int attr_read(hid_t loc, const char *name, hid_t mem_type, void *out)
{
	const hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
	if (a == H5I_INVALID_HID) {
		log_error("attribute %s missing", name);
		return -1;
	}
	const herr_t e = H5Aread(a, mem_type, out);
	H5Aclose(a);
	if (e < 0)
		log_error("attribute %s: read failed", name);
	return e < 0 ? -1 : 0;
}

/* 1 if a link `name` exists under loc, 0 if not, -1 on error. */
static int has_link(hid_t loc, const char *name)
{
	const htri_t t = H5Lexists(loc, name, H5P_DEFAULT);
	return t < 0 ? -1 : t > 0;
}

/* -- file ------------------------------------------------------------------ */

data_id data_open(const char *filename)
{
	const hid_t f = H5Fopen(filename, H5F_ACC_RDWR, H5P_DEFAULT);
	if (f == H5I_INVALID_HID) {
		log_error("data_open(%s): cannot open", filename);
		return DATA_INVALID_FID;
	}
	return (data_id)f;
}

void data_close(data_id fid)
{
	if (fid != DATA_INVALID_FID && fid != DATA_FOLLOWER_FID && fid != 0)
		H5Fclose((hid_t)fid);
}

int data_grp_create(data_id fid, const char *grp_name)
{
	if (fid == DATA_INVALID_FID || fid == 0)
		return -1;
	const hid_t f = (hid_t)fid;
	const int link = has_link(f, grp_name);
	if (link < 0)
		return -1;
	if (link) {
		const htri_t obj = H5Oexists_by_name(f, grp_name, H5P_DEFAULT);
		if (obj > 0)
			return 0;	/* already there */
		/* a soft link to nothing: drop it and create the group */
		if (obj < 0 || H5Ldelete(f, grp_name, H5P_DEFAULT) < 0) {
			log_error("data_grp_create(%s): dangling link", grp_name);
			return -1;
		}
	}
	const hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
	if (lcpl == H5I_INVALID_HID)
		return -1;
	H5Pset_create_intermediate_group(lcpl, 1);
	H5Pset_char_encoding(lcpl, H5T_CSET_UTF8);
	const hid_t g = H5Gcreate2(f, grp_name, lcpl, H5P_DEFAULT, H5P_DEFAULT);
	H5Pclose(lcpl);
	if (g == H5I_INVALID_HID) {
		log_error("data_grp_create(%s): cannot create", grp_name);
		return -1;
	}
	H5Gclose(g);
	return 0;
}

/* -- attributes ------------------------------------------------------------ */

int data_attr_read_dbl(data_id fid, const char *grp_name,
	const char *attr_name, double *a)
{
	const hid_t g = grp_open(fid, grp_name);
	if (g == H5I_INVALID_HID)
		return -1;
	const int rc = attr_read(g, attr_name, H5T_NATIVE_DOUBLE, a);
	H5Gclose(g);
	return rc;
}

static int attr_write(data_id fid, const char *grp_name,
	const char *attr_name, hid_t type, const void *val)
{
	const hid_t g = grp_open(fid, grp_name);
	if (g == H5I_INVALID_HID)
		return -1;
	int rc = -1;
	const hid_t acpl = H5Pcreate(H5P_ATTRIBUTE_CREATE);
	const hid_t sp = H5Screate(H5S_SCALAR);
	if (acpl != H5I_INVALID_HID && sp != H5I_INVALID_HID) {
		H5Pset_char_encoding(acpl, H5T_CSET_UTF8);
		const hid_t a = H5Acreate2(g, attr_name, type, sp, acpl,
			H5P_DEFAULT);
		if (a != H5I_INVALID_HID) {
			rc = H5Awrite(a, type, val) < 0 ? -1 : 0;
			H5Aclose(a);
		}
	}
	if (rc < 0)
		log_error("data_attr_write(%s/%s) failed", grp_name, attr_name);
	if (sp != H5I_INVALID_HID)
		H5Sclose(sp);
	if (acpl != H5I_INVALID_HID)
		H5Pclose(acpl);
	H5Gclose(g);
	return rc;
}

int data_attr_write_ul(data_id fid, const char *grp_name,
	const char *attr_name, unsigned long a)
{
	return attr_write(fid, grp_name, attr_name, H5T_NATIVE_ULONG, &a);
}

int data_attr_write_dbl(data_id fid, const char *grp_name,
	const char *attr_name, double a)
{
	return attr_write(fid, grp_name, attr_name, H5T_NATIVE_DOUBLE, &a);
}

/* -- /state_prep/multidet ---------------------------------------------------- */

int data_muldet_load(data_id fid, struct circ_muldet *md)
{
	memset(md, 0, sizeof *md);
	const hid_t g = grp_open(fid, G_MULTIDET);
	if (g == H5I_INVALID_HID)
		return -1;
	int rc = -1;
	hsize_t dd[2], cd[2];
	double *cf = NULL;
	unsigned char *det = NULL;
	if (dset_dims(g, "dets", 2, dd) < 0 || dset_dims(g, "coeffs", 2, cd) < 0)
		goto out;
	const size_t ndet = dd[0], nqb = dd[1];
	if (cd[0] != ndet || cd[1] != 2 || nqb == 0 || nqb > 64) {
		log_error("data_muldet_load: dets (%zu, %zu), coeffs (%llu, "
			  "%llu)", ndet, nqb, (unsigned long long)cd[0],
			(unsigned long long)cd[1]);
		goto out;
	}
	cf = malloc(sizeof *cf * 2 * (ndet ? ndet : 1));
	det = malloc(ndet * nqb > 0 ? ndet * nqb : 1);
	if (!cf || !det)
		goto out;
	if (dset_read(g, "coeffs", H5T_NATIVE_DOUBLE, cf) < 0 ||
	    dset_read(g, "dets", H5T_NATIVE_UCHAR, det) < 0)
		goto out;
	for (size_t k = 0; k < ndet * nqb; k++)
		if (det[k] > 1) {
			log_error("data_muldet_load: dets[%zu][%zu] = %u is not "
				  "0 or 1", k / nqb, k % nqb, (unsigned)det[k]);
			goto out;
		}
	if (circ_muldet_init(md, ndet) < 0)
		goto out;
	for (size_t d = 0; d < ndet; d++) {
		uint64_t idx = 0;
		for (size_t q = 0; q < nqb; q++)
			idx |= (uint64_t)det[d * nqb + q] << q;
		md->dets[d].idx = idx;
		md->dets[d].cf = CMPLX(cf[2 * d], cf[2 * d + 1]);
	}
	rc = 0;
out:
	free(cf);
	free(det);
	H5Gclose(g);
	return rc;
}

/* -- state-prep dispatch ----------------------------------------------------- */

int data_state_prep_kind(data_id fid, enum stprep_kind *out)
{
	const hid_t g = grp_open(fid, G_STPREP);
	if (g == H5I_INVALID_HID)
		return -1;
	const int md = has_link(g, "multidet"), cm = has_link(g, "coeff_matrix");
	H5Gclose(g);
	if (md < 0 || cm < 0)
		return -1;
	if (md && cm) {
		log_error("simul.h5: both /state_prep/multidet and "
			  "/state_prep/coeff_matrix present");
		return -1;
	}
	if (!md && !cm) {
		log_error("simul.h5: no state-prep subgroup found");
		return -1;
	}
	*out = md ? STPREP_MULTIDET : STPREP_COEFF_MATRIX;
	return 0;
}

/* -- /state_prep/coeff_matrix ------------------------------------------------ */

/* A real (n_sites, n_occ) matrix dataset into a new array. */
static double *read_C(hid_t loc, const char *name, uint32_t n_sites,
	uint32_t n_occ)
{
	hsize_t dims[2];
	if (dset_dims(loc, name, 2, dims) < 0)
		return NULL;
	if (dims[0] != n_sites || dims[1] != n_occ) {
		log_error("%s: shape (%llu, %llu), expected (%u, %u)", name,
			(unsigned long long)dims[0], (unsigned long long)dims[1],
			n_sites, n_occ);
		return NULL;
	}
	const hid_t d = H5Dopen2(loc, name, H5P_DEFAULT);
	const hid_t t = H5Dget_type(d);
	const bool f64 = H5Tget_class(t) == H5T_FLOAT &&
		H5Tget_size(t) == sizeof(double);
	H5Tclose(t);
	H5Dclose(d);
	if (!f64) {
		log_error("%s: not a double dataset", name);
		return NULL;
	}
	const size_t n = (size_t)n_sites * n_occ;
	double *buf = malloc(sizeof *buf * (n ? n : 1));
	if (buf && dset_read(loc, name, H5T_NATIVE_DOUBLE, buf) < 0) {
		free(buf);
		buf = NULL;
	}
	return buf;
}

int data_coeff_matrix_load(data_id fid, struct data_coeff_matrix *cm)
{
	memset(cm, 0, sizeof *cm);
	const hid_t g = grp_open(fid, G_COEFFMAT);
	if (g == H5I_INVALID_HID)
		return -1;
	int rc = -1;
	uint8_t cs = 0, tap = 0;
	if (attr_read(g, "n_qubits", H5T_NATIVE_UINT32, &cm->nqb) < 0 ||
	    attr_read(g, "n_sites", H5T_NATIVE_UINT32, &cm->n_sites) < 0 ||
	    attr_read(g, "n_alpha", H5T_NATIVE_UINT32, &cm->n_alpha) < 0 ||
	    attr_read(g, "n_beta", H5T_NATIVE_UINT32, &cm->n_beta) < 0 ||
	    attr_read(g, "closed_shell", H5T_NATIVE_UINT8, &cs) < 0 ||
	    attr_read(g, "tapered", H5T_NATIVE_UINT8, &tap) < 0)
		goto out;
	cm->closed_shell = cs != 0;
	cm->tapered = tap != 0;
	const int csf = has_link(g, "csf");
	if (csf < 0)
		goto out;
	if (!csf) {
		cm->C_alpha = read_C(g, "C_alpha", cm->n_sites, cm->n_alpha);
		if (!cm->C_alpha)
			goto out;
		if (!cm->closed_shell) {
			cm->C_beta = read_C(g, "C_beta", cm->n_sites,
				cm->n_beta);
			if (!cm->C_beta)
				goto out;
		}
		rc = 0;
		goto out;
	}
	const hid_t c = H5Gopen2(g, "csf", H5P_DEFAULT);
	uint32_t ncomp = 0;
	const int ok = c != H5I_INVALID_HID &&
		attr_read(c, "n_components", H5T_NATIVE_UINT32, &ncomp) == 0;
	if (c != H5I_INVALID_HID)
		H5Gclose(c);
	if (!ok)
		goto out;
	if (ncomp == 0) {
		log_error("data_coeff_matrix_load: csf/ with no components");
		goto out;
	}
	cm->blocks = calloc(ncomp, sizeof *cm->blocks);
	if (!cm->blocks)
		goto out;
	cm->n_components = ncomp;
	for (uint32_t k = 0; k < ncomp; k++) {
		char sub[32];
		snprintf(sub, sizeof sub, "csf/%u", k);
		const hid_t b = H5Gopen2(g, sub, H5P_DEFAULT);
		if (b == H5I_INVALID_HID) {
			log_error("data_coeff_matrix_load: %s missing", sub);
			goto out;
		}
		struct data_coeff_block *B = &cm->blocks[k];
		int bad = attr_read(b, "coefficient", H5T_NATIVE_DOUBLE,
			&B->cf) < 0;
		if (!bad)
			bad = !(B->C_alpha = read_C(b, "C_alpha", cm->n_sites,
					cm->n_alpha));
		if (!bad && !cm->closed_shell)
			bad = !(B->C_beta = read_C(b, "C_beta", cm->n_sites,
					cm->n_beta));
		H5Gclose(b);
		if (bad)
			goto out;
	}
	rc = 0;
out:
	H5Gclose(g);
	if (rc < 0)
		data_coeff_matrix_free(cm);
	return rc;
}

/* -- /pauli_hamil ------------------------------------------------------------ */

int data_hamil_load(data_id fid, struct circ_hamil *hm)
{
	const hid_t g = grp_open(fid, G_HAMIL);
	if (g == H5I_INVALID_HID)
		return -1;
	int rc = -1;
	double norm = 0.0, *cf = NULL;
	unsigned char *pb = NULL;
	hsize_t pd[2], cd[1];
	if (dset_dims(g, "paulis", 2, pd) < 0 || dset_dims(g, "coeffs", 1, cd) < 0
	    || attr_read(g, "normalization", H5T_NATIVE_DOUBLE, &norm) < 0)
		goto out;
	const size_t nterm = pd[0], nqb = pd[1];
	if (cd[0] != nterm || nqb == 0 || nqb > 64) {
		log_error("data_hamil_load: paulis (%zu, %zu), coeffs (%llu)",
			nterm, nqb, (unsigned long long)cd[0]);
		goto out;
	}
	cf = malloc(sizeof *cf * (nterm ? nterm : 1));
	pb = malloc(nterm * nqb > 0 ? nterm * nqb : 1);
	if (!cf || !pb)
		goto out;
	if (dset_read(g, "coeffs", H5T_NATIVE_DOUBLE, cf) < 0 ||
	    dset_read(g, "paulis", H5T_NATIVE_UCHAR, pb) < 0)
		goto out;
	for (size_t k = 0; k < nterm * nqb; k++)
		if (pb[k] > 3) {
			log_error("data_hamil_load: paulis[%zu][%zu] = %u is not "
				  "I/X/Y/Z", k / nqb, k % nqb, (unsigned)pb[k]);
			goto out;
		}
	if (circ_hamil_init(hm, (uint32_t)nqb, nterm) < 0)
		goto out;
	for (size_t t = 0; t < nterm; t++) {
		struct paulis op = paulis_new();
		for (size_t q = 0; q < nqb; q++)
			paulis_set(&op, (enum pauli_op)pb[t * nqb + q],
				(uint32_t)q);
		hm->terms[t].op = op;
		hm->terms[t].cf = cf[t] * norm;
	}
	rc = 0;
out:
	free(cf);
	free(pb);
	H5Gclose(g);
	return rc;
}

/* -- per-step writer ----------------------------------------------------------- */

int data_circ_writer_init(data_id fid, const char *grp_name, size_t n_steps,
	struct data_circ_writer *w)
{
	memset(w, 0, sizeof *w);
	if (fid == 0)
		return 0;	/* disabled: writes do nothing */
	if (data_grp_create(fid, grp_name) < 0)
		return -1;
	const hid_t g = grp_open(fid, grp_name);
	if (g == H5I_INVALID_HID)
		return -1;
	hid_t d = H5I_INVALID_HID;
	const int exists = has_link(g, D_VALUES);
	if (exists > 0) {
		/* a values dataset from an earlier run: written over */
		d = H5Dopen2(g, D_VALUES, H5P_DEFAULT);
	} else if (exists == 0) {
		const hsize_t dims[2] = { n_steps, 2 };
		const hid_t sp = H5Screate_simple(2, dims, NULL);
		const hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
		const double nanv = nan("");
		if (sp != H5I_INVALID_HID && dcpl != H5I_INVALID_HID &&
		    H5Pset_fill_value(dcpl, H5T_NATIVE_DOUBLE, &nanv) >= 0 &&
		    H5Pset_alloc_time(dcpl, H5D_ALLOC_TIME_EARLY) >= 0 &&
		    H5Pset_fill_time(dcpl, H5D_FILL_TIME_ALLOC) >= 0)
			d = H5Dcreate2(g, D_VALUES, H5T_IEEE_F64LE, sp,
				H5P_DEFAULT, dcpl, H5P_DEFAULT);
		if (dcpl != H5I_INVALID_HID)
			H5Pclose(dcpl);
		if (sp != H5I_INVALID_HID)
			H5Sclose(sp);
		if (d != H5I_INVALID_HID)
			H5Fflush((hid_t)fid, H5F_SCOPE_GLOBAL);
	}
	H5Gclose(g);
	if (d == H5I_INVALID_HID) {
		log_error("data_circ_writer_init(%s): no values dataset",
			grp_name);
		return -1;
	}
	const hsize_t row[2] = { 1, 2 };
	const hid_t ms = H5Screate_simple(2, row, NULL);
	if (ms == H5I_INVALID_HID) {
		H5Dclose(d);
		return -1;
	}
	w->fid = fid;
	w->dset = (int64_t)d;
	w->mspace = (int64_t)ms;
	w->n_steps = n_steps;
	return 0;
}

int data_circ_write_step(struct data_circ_writer *w, size_t step_idx,
	_Complex double z)
{
	if (w->fid == 0)
		return 0;
	if (step_idx >= w->n_steps) {
		log_error("data_circ_write_step: step %zu of %zu", step_idx,
			w->n_steps);
		return -1;
	}
	const hid_t d = (hid_t)w->dset;
	const hid_t fs = H5Dget_space(d);
	const hsize_t start[2] = { step_idx, 0 }, count[2] = { 1, 2 };
	const double v[2] = { creal(z), cimag(z) };
	int rc = -1;
	if (fs != H5I_INVALID_HID &&
	    H5Sselect_hyperslab(fs, H5S_SELECT_SET, start, NULL, count,
		    NULL) >= 0 &&
	    H5Dwrite(d, H5T_NATIVE_DOUBLE, (hid_t)w->mspace, fs, H5P_DEFAULT,
		    v) >= 0 &&
	    H5Fflush((hid_t)w->fid, H5F_SCOPE_GLOBAL) >= 0)
		rc = 0;
	if (fs != H5I_INVALID_HID)
		H5Sclose(fs);
	if (rc < 0)
		log_error("data_circ_write_step: step %zu failed", step_idx);
	return rc;
}

void data_circ_writer_close(struct data_circ_writer *w)
{
	if (!w || w->fid == 0)
		return;
	if (w->dset > 0)
		H5Dclose((hid_t)w->dset);
	if (w->mspace > 0)
		H5Sclose((hid_t)w->mspace);
	memset(w, 0, sizeof *w);
}

/* -- read-back for callers without HDF5 bindings ------------------------------- */

/* The /<grp>/values rows of a simul.h5 file: up to cap (re, im) pairs into
 * out, the row count into *nrows.  0 or -1.  No counterpart in the
 * reference (tests and tools read results through it). */
int ph2_h5_read_values(const char *path, const char *grp, double *out,
	size_t cap, size_t *nrows)
{
	const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
	if (f == H5I_INVALID_HID)
		return -1;
	int rc = -1;
	char name[256];
	snprintf(name, sizeof name, "%s/%s", grp, D_VALUES);
	hsize_t dims[2];
	if (dset_dims(f, name, 2, dims) == 0 && dims[1] == 2) {
		*nrows = dims[0];
		if (dims[0] <= cap)
			rc = dset_read(f, name, H5T_NATIVE_DOUBLE, out);
	}
	H5Fclose(f);
	return rc;
}
