#ifndef PHASE2_DATA_H
#define PHASE2_DATA_H

/*
 * The simul.h5 data layer of ph2run (replaces include/ph2run/data.h and
 * ph2run/data.c of the reference; same names, types, macros, return values
 * and file layout, so the reference's ph2run.c builds against it
 * unchanged).  Implemented in phase2_amd/csrc/h5/data_h5.c as
 * libphase2_h5.so over the HDF5 C library.  One GPU, one process: the
 * process that opens the file does all I/O (the reference's followers that
 * receive broadcasts do not exist here), so DATA_FOLLOWER_FID is never
 * returned.  File layout: doc/simul-h5-specs.md of the reference.
 */

#include <stddef.h>
#include <stdint.h>

struct circ_hamil;

#define DATA_INVALID_FID INT64_C(-1)
/* A handle of a follower rank in the reference's MPI layer; kept for
 * source compatibility, never produced here. */
#define DATA_FOLLOWER_FID INT64_C(-2)

/* Group and attribute names the circuit drivers write per run. */
#define DATA_CIRCTROTT "circ_trott"
#define DATA_CIRCTROTT_DELTA "delta"

#define DATA_CIRCTROTT2 "circ_trott2"
#define DATA_CIRCTROTT2_DELTA "delta"

#define DATA_CIRCQDRIFT "circ_qdrift"
#define DATA_CIRCQDRIFT_DEPTH "depth"
#define DATA_CIRCQDRIFT_NUMSAMPLES "num_samples"
#define DATA_CIRCQDRIFT_STEPSIZE "step_size"
#define DATA_CIRCQDRIFT_SEED "seed"

#define DATA_CIRCCMPSIT "circ_cmpsit"
#define DATA_CIRCCMPSIT_DEPTH "depth"
#define DATA_CIRCCMPSIT_LENGTH "length"
#define DATA_CIRCCMPSIT_ANGLEDET "angle_det"
#define DATA_CIRCCMPSIT_ANGLERAND "angle_rand"
#define DATA_CIRCCMPSIT_STEPS "steps"
#define DATA_CIRCCMPSIT_SEED "seed"

/* An open simul.h5 (an HDF5 file id). */
typedef int64_t data_id;

/* Open read-write; DATA_INVALID_FID on failure. */
data_id data_open(const char *filename);
void data_close(data_id);

/* Create a group (and missing parents); an existing group is success, a
 * dangling soft link of that name is removed first.  0 or -1. */
int data_grp_create(data_id fid, const char *grp_name);

int data_attr_read_dbl(data_id fid, const char *grp_name,
	const char *attr_name, double *a);

#define data_attr_read(fid, grp_name, attr_name, attr_buf)                     \
	_Generic((attr_buf),                                                   \
		double *: data_attr_read_dbl)(                                 \
		fid, grp_name, attr_name, attr_buf)

/* Create a scalar attribute holding a (fails if it exists).  0 or -1. */
int data_attr_write_ul(data_id fid, const char *grp_name,
	const char *attr_name, unsigned long a);
int data_attr_write_dbl(data_id fid, const char *grp_name,
	const char *attr_name, double a);

#define data_attr_write(fid, grp_name, attr_name, attr)                        \
	_Generic((attr),                                                       \
		double: data_attr_write_dbl,                                   \
		unsigned long: data_attr_write_ul)(                            \
		fid, grp_name, attr_name, attr)

/* /state_prep/multidet into { idx, cf } rows (qubit j of a determinant at
 * bit j); occupation bytes other than 0 and 1 are rejected.  Release with
 * circ_muldet_free.  0 or -1. */
struct circ_muldet;
int data_muldet_load(data_id fid, struct circ_muldet *md);

/* Which of /state_prep/multidet and /state_prep/coeff_matrix the file
 * holds: exactly one must be present (else -1, *out unchanged). */
enum stprep_kind;
struct data_coeff_matrix;
int data_state_prep_kind(data_id fid, enum stprep_kind *out);

/* /state_prep/coeff_matrix: attributes, C_alpha (and C_beta unless
 * closed_shell) of shape (n_sites, n_occ), or the csf/ components
 * (n_components >= 1, each with its coefficient).  Release with
 * data_coeff_matrix_free.  0 or -1. */
int data_coeff_matrix_load(data_id fid, struct data_coeff_matrix *cm);

/* /pauli_hamil: coefficients times the normalization attribute, Pauli
 * bytes I=0, X=1, Y=2, Z=3 (others rejected).  Release with
 * circ_hamil_free.  0 or -1. */
int data_hamil_load(data_id fid, struct circ_hamil *hm);

/* Per-step output /<grp>/values, shape (n_steps, 2), NaN until written;
 * each write is flushed to the file.  fid == 0 makes a disabled writer
 * whose writes do nothing. */
struct data_circ_writer {
	data_id fid;
	int64_t dset;	/* the open values dataset */
	int64_t mspace;	/* (1, 2) memory dataspace */
	size_t n_steps;
};

int data_circ_writer_init(data_id fid, const char *grp_name, size_t n_steps,
	struct data_circ_writer *w);
int data_circ_write_step(struct data_circ_writer *w, size_t step_idx,
	_Complex double z);
void data_circ_writer_close(struct data_circ_writer *w);

#endif // PHASE2_DATA_H
