/*
 * ifos2d_dropin.h -- the reference's hot-path entry points, same names and
 * argument lists as IFOS2D 2.0.3's src/fd.h, implemented on the MI355X by
 * libifos2d_hip.so.
 *
 * Wavefield, material, memory-variable and CPML-memory arrays must be device
 * arrays from ifos_dev_matrix()/ifos_dev_f3tensor() (same index bounds as the
 * reference's matrix()/f3tensor() calls, src/IFOS2D.c:371-641).  The CPML
 * profile vectors, hc, bip..cjm, etajm/peta, srcpos_loc, signals and the
 * seismogram sections stay host arrays as in the reference.  Calls are queued
 * in order on one device stream and return without waiting for the device;
 * call ifos_dev_sync() before the host reads device data or the seismogram
 * sections (seismo_ssg gathers the samples on the device; ifos_dev_sync()
 * writes them into the host sections) and use ifos_dev_read() for single
 * values (the stability probe of src/IFOS2D.c:1660).  A device tensor keeps its elements plane by plane
 * (third index slowest), so every kernel reads and writes whole planes; the
 * float*** it returns is a handle for these calls, not for host indexing,
 * and ifos_dev_copy_array() converts from and to util.c's [j][i][l] block.
 * update_v_PML and psource keep a device copy of the whole signals block
 * and of the source positions between calls; each call compares the
 * samples and positions it reads with that copy and uploads them again when
 * they differ, so changed host values take effect at the next step as in
 * the reference.  The layer serves one host thread, like the reference's
 * driver (one process per MPI rank).
 *
 * Globals read by these functions, by the reference names (src/globvar.h):
 * DT, DH, NX, NY, NT, FDORDER, FW, FREE_SURF, BOUNDARY, L, FL, F_REF,
 * PARAMETERIZATION (1), VELOCITY, ADJOINT_TYPE, SOURCE_TYPE (written by
 * update_v_PML as in src/update_v_PML.c:558), NPROCX = NPROCY = 1, SEISMO
 * (1 or 2).  The library defines them; a driver that defines them itself
 * overrides them.
 */
#ifndef IFOS2D_DROPIN_H
#define IFOS2D_DROPIN_H

#ifdef __cplusplus
extern "C" {
#endif

/* device Numerical-Recipes arrays (shapes of src/util.c:95,184,279) */
float **ifos_dev_matrix(int nrl, int nrh, int ncl, int nch);
float ***ifos_dev_f3tensor(int nrl, int nrh, int ncl, int nch, int ndl, int ndh);
float *ifos_dev_vector(int nl, int nh);
void ifos_dev_free_array(void *arr);
/* copy a whole device array from (dir 0) or to (dir 1) host memory laid out
 * like the reference allocation: host_first_elem = &m[nrl][ncl] */
int ifos_dev_copy_array(void *dev_arr, const void *host_first_elem, int dir);
int ifos_dev_sync(void);
float ifos_dev_read(float **m, int j, int i);

/* replaces update_v_PML()            src/update_v_PML.c:29, src/fd.h:384-389 */
void update_v_PML(int nx1, int nx2, int ny1, int ny2, int nt, float **vx, float **vxp1, float **vxm1, float **vy,
                  float **vyp1, float **vym1, float **uttx, float **utty, float **sxx, float **syy, float **sxy,
                  float **rip, float **rjp, float **srcpos_loc, float **signals, float **signals1, int nsrc,
                  float **absorb_coeff, float *hc, int infoout, int sw, float *K_x, float *a_x, float *b_x,
                  float *K_x_half, float *a_x_half, float *b_x_half, float *K_y, float *a_y, float *b_y,
                  float *K_y_half, float *a_y_half, float *b_y_half, float **psi_sxx_x, float **psi_syy_y,
                  float **psi_sxy_y, float **psi_sxy_x);

/* replaces update_s_elastic_PML()    src/update_s_elastic_PML.c:28, src/fd.h:352-357 */
void update_s_elastic_PML(int nx1, int nx2, int ny1, int ny2, float **vx, float **vy, float **ux, float **uy,
                          float **uxy, float **uyx, float **sxx, float **syy, float **sxy, float **pi, float **u,
                          float **uipjp, float **absorb_coeff, float **rho, float *hc, int infoout, float *K_x,
                          float *a_x, float *b_x, float *K_x_half, float *a_x_half, float *b_x_half, float *K_y,
                          float *a_y, float *b_y, float *K_y_half, float *a_y_half, float *b_y_half,
                          float **psi_vxx, float **psi_vyy, float **psi_vxy, float **psi_vyx);

/* replaces update_s_visc_PML()       src/update_s_visc_PML.c:28, src/fd.h:363-369 */
void update_s_visc_PML(int nx1, int nx2, int ny1, int ny2, float **vx, float **vy, float **ux, float **uy,
                       float **uxy, float **uyx, float **sxx, float **syy, float **sxy, float **pi, float **u,
                       float **uipjp, float **rho, float *hc, int infoout, float ***r, float ***p, float ***q,
                       float **fipjp, float **f, float **g, float *bip, float *bjm, float *cip, float *cjm,
                       float ***d, float ***e, float ***dip, float *K_x, float *a_x, float *b_x, float *K_x_half,
                       float *a_x_half, float *b_x_half, float *K_y, float *a_y, float *b_y, float *K_y_half,
                       float *a_y_half, float *b_y_half, float **psi_vxx, float **psi_vyy, float **psi_vxy,
                       float **psi_vyx);

/* replaces psource()                 src/psource.c:26, src/fd.h:206-207 */
void psource(int nt, float **sxx, float **syy, float **sp, float **srcpos_loc, float **signals, int nsrc, int sw);

/* replaces surface_elastic_PML()     src/surface_elastic_PML.c:26, src/fd.h:286 */
void surface_elastic_PML(int ndepth, float **vx, float **vy, float **sxx, float **syy, float **sxy, float **syz,
                         float **pi, float **u, float **rho, float *hc, float *K_x, float *a_x, float *b_x,
                         float **psi_vxx, float **ux, float **uy, float **uxy, float **uyz, float **sxz,
                         float **uxz);

/* replaces surface_PML()             src/surface_PML.c:25, src/fd.h:288 */
void surface_PML(int ndepth, float **vx, float **vy, float **sxx, float **syy, float **sxy, float **syz,
                 float ***p, float ***q, float **ppi, float **pu, float **prho, float **ptaup, float **ptaus,
                 float *etajm, float *peta, float *hc, float *K_x, float *a_x, float *b_x, float **psi_vxx,
                 float **ux, float **uy, float **uxy, float **uyz, float **sxz, float **uxz);

/* replaces seismo_ssg()              src/seismo_ssg.c:26, src/fd.h:262-265 (SEISMO 1, 2);
 * the samples reach the host sections at the next ifos_dev_sync() */
void seismo_ssg(int lsamp, int ntr, int **recpos, float **sectionvx, float **sectionvy, float **sectionvz,
                float **sectionp, float **sectioncurl, float **sectiondiv, float **vx, float **vy, float **vz,
                float **sxx, float **syy, float **sp, float **pi, float **u, float *hc);

/* replace exchange_v() / exchange_s() src/exchange_v.c:28, src/exchange_s.c:28, src/fd.h:116-124,
 * on the one rank that holds the whole grid: with BOUNDARY 1 the halo
 * columns take the opposite edge (periodic sides), with BOUNDARY 0 nothing
 * changes.  The buffers and the request pointers are not used;
 * wavetyp_start must be 1.  After mpi.h the prototypes are src/fd.h's own
 * (MPI_Request *), so a driver may include both headers; without MPI the
 * request pointers are plain pointers (the same calling convention). */
#ifdef MPI_VERSION
#define IFOS_MPI_REQUEST MPI_Request
#else
#define IFOS_MPI_REQUEST void
#endif
void exchange_v(float **vx, float **vy, float **vz, float **bufferlef_to_rig, float **bufferrig_to_lef,
                float **buffertop_to_bot, float **bufferbot_to_top, IFOS_MPI_REQUEST *req_send,
                IFOS_MPI_REQUEST *req_rec, int wavetyp_start);
void exchange_s(float **sxx, float **syy, float **sxy, float **sxz, float **syz, float **bufferlef_to_rig,
                float **bufferrig_to_lef, float **buffertop_to_bot, float **bufferbot_to_top,
                IFOS_MPI_REQUEST *req_send, IFOS_MPI_REQUEST *req_rec, int wavetyp_start);
#undef IFOS_MPI_REQUEST

/* replaces zero_fdveps()             src/zero_fdveps.c:28, src/fd.h:428-429 */
void zero_fdveps(int ny1, int ny2, int nx1, int nx2, float **vx, float **vy, float **vz, float **sxx, float **syy,
                 float **sxy, float **sxz, float **syz, float **vxm1, float **vym1, float **uxy, float **vxp1,
                 float **vyp1, float **psi_sxx_x, float **psi_sxy_x, float **psi_sxz_x, float **psi_vxx,
                 float **psi_vyx, float **psi_vzx, float **psi_syy_y, float **psi_sxy_y, float **psi_syz_y,
                 float **psi_vyy, float **psi_vxy, float **psi_vzy, float **psi_vxxs);

/* replaces zero_fdveps_visc()        src/zero_fdveps_visc.c:28, src/fd.h:430-433 */
void zero_fdveps_visc(int ny1, int ny2, int nx1, int nx2, float **vx, float **vy, float **vz, float **sxx,
                      float **syy, float **sxy, float **sxz, float **syz, float **vxm1, float **vym1, float **uxy,
                      float **vxp1, float **vyp1, float **psi_sxx_x, float **psi_sxy_x, float **psi_sxz_x,
                      float **psi_vxx, float **psi_vyx, float **psi_vzx, float **psi_syy_y, float **psi_sxy_y,
                      float **psi_syz_y, float **psi_vyy, float **psi_vxy, float **psi_vzy, float **psi_vxxs,
                      float ***pr, float ***pp, float ***pq, float ***pt, float ***po);

/* The loops main runs inline, promoted to entry points on device arrays
 * (WAVETYPE 1, not ACOUSTIC, IDXI = IDYI = IDX = IDY = 1).  The snapshot
 * tensors are f3tensor(-nd+1,NY+nd,-nd+1,NX+nd,1,NT/DTINV) from
 * ifos_dev_f3tensor(), time fastest as in the reference. */

/* replaces the snapshot store of src/IFOS2D.c:1805-1871 (slot hin): vxp1,
 * vyp1 and, with VELOCITY 0, sxx, syy, sxy, else ux, uy, uxy */
void ifos_snapshot_store(int hin, float **pvxp1, float **pvyp1, float **psxx, float **psyy, float **psxy,
                         float **ux, float **uy, float **uxy, float ***forward_prop_rho_x,
                         float ***forward_prop_rho_y, float ***forward_prop_x, float ***forward_prop_y,
                         float ***forward_prop_u);

/* replaces the zero-lag correlation of src/IFOS2D.c:2381-2419 with snapshot
 * slot k = NTDTINV-hin+1, accumulated into the waveconv_*_shot arrays */
void ifos_correlate(int k, float **pvxp1, float **pvyp1, float **psxx, float **psyy, float **psxy, float **prho,
                    float **ppi, float **pu, float ***forward_prop_rho_x, float ***forward_prop_rho_y,
                    float ***forward_prop_x, float ***forward_prop_y, float ***forward_prop_u,
                    float **waveconv_shot, float **waveconv_u_shot, float **waveconv_rho_shot);

/* replaces the chain rule of src/IFOS2D.c:2471-2611 (PARAMETERIZATION 1):
 * waveconv_shot, waveconv_u_shot and waveconv_rho_shot become the Vp, Vs
 * and rho gradients of the shot; waveconv_lam, waveconv_mu and
 * waveconv_rho_s (each may be NULL) receive lambda', mu' and rho' */
void ifos_gradient_assemble(float **prho, float **ppi, float **pu, float **waveconv_shot, float **waveconv_u_shot,
                            float **waveconv_rho_shot, float **waveconv_lam, float **waveconv_mu,
                            float **waveconv_rho_s);

#ifdef __cplusplus
}
#endif
#endif
