/*
 * ifos2d_formats.h -- the reference's file formats on the P-SV path, so a
 * run of this library reads and writes what an IFOS2D run does:
 *   - the quasi-JSON parameter file (src/json_parser.c, src/read_par_json.c),
 *   - binary model files (src/readmod_elastic.c, src/writemod.c),
 *   - SU seismograms (src/outseis_glob.c, SEIS_FORMAT 1) and their reading
 *     back as observed data (src/inseis.c).
 * Host-side C, plain pointers.  0 on success, negative on failure with the
 * message in ifos_last_error().
 */
#ifndef IFOS2D_FORMATS_H
#define IFOS2D_FORMATS_H

#include "ifos2d_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ifos_json ifos_json;

/* replaces read_objects_from_intputfile() src/json_parser.c:26: the
 * "name" : "value" objects of the file, up to five per line
 * ("A, B" : "1, 2"), lines naming "comment"/"Comment" skipped, names and
 * values cut to their first blank-delimited word. */
int  ifos_json_load(const char *path, ifos_json **out);
void ifos_json_free(ifos_json *j);
int  ifos_json_count(const ifos_json *j);
/* name and value of object k (file order), or NULL */
const char *ifos_json_name(const ifos_json *j, int k);
const char *ifos_json_value(const ifos_json *j, int k);
/* the value of the first object called `name`, or NULL */
const char *ifos_json_get(const ifos_json *j, const char *name);
/* numeric values as get_float_from_objectlist / get_int_from_objectlist
 * (src/json_parser.c:264-385) read them: 0 found, 1 absent, -2 malformed */
int  ifos_json_float(const ifos_json *j, const char *name, float *out);
int  ifos_json_int(const ifos_json *j, const char *name, int *out);

/* The run parameters of the P-SV path from a parameter file, as
 * read_par_json (src/read_par_json.c) sets the globals: NX, NY, DH, DT,
 * NT = iround(TIME/DT) (src/IFOS2D.c:257), FDORDER, MAXRELERROR, FW,
 * FREE_SURF, BOUNDARY, VPPML, FPML, npower, k_max_PML, L, FL1..FL5, TAU,
 * F_REF, VELOCITY, ADJOINT_TYPE, PARAMETERIZATION, DTINV, SOURCE_SHAPE,
 * SOURCE_TYPE, TRKILL.  Absent optional keys keep the values *c holds. */
int ifos_config_from_json(const char *path, ifos_config *c);

/* Binary model files: NX*NY native floats written x-outer, y-inner
 * (src/writemod.c:36-41); arrays here are [NY][NX], x fastest. */
int ifos_model_read(const char *path, int NX, int NY, float *m);
int ifos_model_write(const char *path, int NX, int NY, const float *m);

/* SU seismograms as outseis_glob writes them for one shot (SEIS_FORMAT 1):
 * per trace a 240-byte SEG-Y trace header (tracl, cdp, trid, offset,
 * gelev, sdepth, swdep, scalel/scalco = -1000, sx, gx, ns, dt, d1 as there)
 * followed by ns floats.  data [ntr][ns]; recpos [ntr][2] grid indices;
 * xs, ys: source position in metres; ndt: the output sample stride (NDT). */
int ifos_su_write(const char *path, int ntr, int ns, const float *data, const int *recpos, float DH, float DT,
                  int ndt, float xs, float ys, float refrec_x, float refrec_y);
/* Read an SU file's traces: ntr_out traces of ns samples each into data
 * [max_tr][ns] (as inseis reads observed data); fails if the file's
 * sample count differs from ns. */
int ifos_su_read(const char *path, int max_tr, int ns, float *data, int *ntr_out);

/* replaces inseis() src/inseis.c:24 (SEIS_FORMAT 1): the ntr traces of an SU
 * file as observed data -- sample 1 of every trace is set to zero, samples
 * 2..ns are the file's (src/inseis.c:75-83); fails unless the file holds
 * exactly ntr traces of ns samples. */
int ifos_su_read_observed(const char *path, int ntr, int ns, float *data);

/* replaces sources() src/sources.c:27 with SRCREC 1: SOURCE_FILE's first
 * line holds the count; each further line x, z, y, tshift, fc, amp
 * [, azimuth [, type]] (z dropped, azimuth 0 and type SOURCE_TYPE when
 * absent; fewer than six values fail as there).  srcpos [max_src][8] rows
 * x (m), y (m), 0, tshift, fc, amp, azimuth, type; *fc_max = the largest fc
 * (the reference's TS = 1/fc_max). */
int ifos_sources_read(const char *path, int SOURCE_TYPE, int max_src, float *srcpos, int *nsrc_out, float *fc_max);
/* replaces splitsrc() src/splitsrc.c:25 on one rank: grid indices
 * iround(x/DH), iround(y/DH) in the rows' first two columns (srcpos rows of
 * IFOS_SRC_STRIDE, the form ifos_engine_set_shot takes); sources whose
 * indices fall outside the rank's truncating-division test are dropped. */
int ifos_sources_to_grid(int nsrc, const float *srcpos_m, float DH, int NX, int NY, float *srcpos, int *nsrc_out);
/* replaces receiver() src/receiver.c:27 with READREC 1: one "x y" pair
 * (metres) per line of REC_FILE, grid indices iround((x + REFREC[1])/DH),
 * iround((y + REFREC[2])/DH); a receiver on the grid point of an earlier
 * one is dropped.  recpos [max_tr][2]. */
int ifos_receivers_read(const char *path, float DH, float refrec_x, float refrec_y, int max_tr, int *recpos,
                        int *ntr_out);
/* receiver() with READREC 0: the straight line from (XREC1, YREC1) to
 * (XREC2, YREC2) every NGEOPH grid points, horizontal unless steeper than
 * 45 degrees, with the reference's integer slope (src/receiver.c:159-204). */
int ifos_receivers_line(float DH, float XREC1, float YREC1, float XREC2, float YREC2, int NGEOPH, int max_tr,
                        int *recpos, int *ntr_out);

/* replaces write_matrix_disk() src/matrix_operations.c:24 on one rank: the
 * gradient or model m [NY][NX] to "<path_name>.bin", x-outer (the model
 * file layout; the reference's per-rank file merged by mergemod). */
int ifos_matrix_dump(const char *path_name, int NX, int NY, const float *m);

/* Trace kill tables [ntr_glob][nsrc_glob] (1 = trace not used), as
 * calc_res builds kill_tmp (src/calc_res.c:66-126).  A shot's kill mask for
 * ifos_engine_residual is its column: kill[r] = table[r][shot].
 * ifos_trkill_read: the integers of TRKILL_FILE.dat, receiver-major (fails,
 * as the reference does, if the file ends while a value is read, so the
 * last value must be followed by white space).  ifos_trkill_offsets:
 * create_trkill_table (src/create_trkill_table.c:21-43) -- sets 1 where the
 * source-receiver offset lies strictly between lower and upper; srcpos
 * [nsrc_glob][2] in metres, recpos [ntr_glob][2] grid indices; entries
 * outside the band are left as they are (TRKILL_OFFSET 2 adds to a file). */
int ifos_trkill_read(const char *path, int ntr_glob, int nsrc_glob, int *table);
int ifos_trkill_offsets(int ntr_glob, const int *recpos, int nsrc_glob, const float *srcpos, float DH, float lower,
                        float upper, int *table);

#ifdef __cplusplus
}
#endif
#endif
