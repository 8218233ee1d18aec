/*
 * cloudsc_field.h -- device-resident blocked fields for the cloud scheme, with the ownership model of the IFS
 * field API, and the physics step that runs on them.
 *
 * Reference interfaces replaced:
 *   ifs-source/contrib/field_api/field_RANKSUFF_module.fypp   FIELD_nRB owner / wrapper objects:
 *     status bits NDEVFRESH, NHSTFRESH, UNALLOCATED, UNINITIALIZED (:25-28), WRAP (:106-131),
 *     OWNER_INIT (:133-183, delayed allocation and INIT_VALUE), GET_VIEW (:193-220), DELETE_DEVICE (:222-229),
 *     WRAPPER_FINAL (:238-244, copies device data back to the wrapped array), GET_HOST_DATA (:377-396),
 *     OWNER_GET_HOST_DATA (:398-413), GET_DEVICE_DATA (:456-478), OWNER_GET_DEVICE_DATA (:480-497),
 *     the _RDONLY / _RDWR and SYNC_HOST / SYNC_DEVICE variants (:415-449 and after :497)
 *   ifs-source/contrib/field_api/field_factory_module.fypp    FIELD_NEW (owner, wrapper), FIELD_DELETE
 *   ifs-source/contrib/field_api/field_basic_module.F90       GPU_STATS transfer counters (:19-27)
 * The field API moves data with OpenACC; on gfx950 the device copy is HIP memory, host<->device copies are
 * hipMemcpyAsync on the caller's stream (field API's QUEUE), and owners allocate page-locked host memory.
 *
 * Semantics, as in the field API:
 *   - a wrapper starts HSTFRESH; an owner starts UNINITIALIZED (both fresh: nothing to copy) or, with
 *     `delayed`, UNALLOCATED until first use; an owner with an init value starts HSTFRESH with every element set;
 *   - get_device_data copies host->device only when the device copy is not fresh, then marks it fresh; with WR in
 *     `mode` it marks the host copy stale.  get_host_data is the mirror image;
 *   - get_view(block) returns the host slice of one last-dimension index; it is an error unless the host copy is
 *     fresh, and it marks the device copy stale (the view may be written);
 *   - delete of a wrapper first makes the wrapped host array fresh (device->host copy if needed).
 * Only persistent fields (the last dimension is the block index NGPBLKS) are supported; the field API's
 * per-thread temporaries (PERSISTENT = .FALSE.) are host-only scratch and stay on the host side.
 *
 * Calls on one field are not thread-safe (as in the field API); different fields may be used concurrently.
 */
#ifndef CLOUDSC_FIELD_H
#define CLOUDSC_FIELD_H

#include <stddef.h>
#include <stdint.h>

#include "cloudsc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cloudsc_field cloudsc_field;

/* Access modes (field API NRD, NWR) and status bits (NDEVFRESH, NHSTFRESH, UNALLOCATED). */
#define CLOUDSC_FIELD_RD 1
#define CLOUDSC_FIELD_WR 2
#define CLOUDSC_FIELD_RDWR 3
#define CLOUDSC_FIELD_DEVFRESH 1
#define CLOUDSC_FIELD_HSTFRESH 2
#define CLOUDSC_FIELD_UNALLOCATED 4

/* Wrapper around an existing host array of `rank` (1..5) dimensions, Fortran order, extents ubounds-lbounds+1.
 * (FIELD_NEW(..., DATA=...)). */
int cloudsc_field_wrap(void *host, size_t elem_size, int rank, const int64_t *lbounds, const int64_t *ubounds,
                       cloudsc_field **out);
/* Owner with library-allocated page-locked host memory (FIELD_NEW(..., UBOUNDS=..., PERSISTENT=.TRUE.)).
 * lbounds may be NULL (all 1); delayed: allocate on first use; init_value: NULL or one element to fill with. */
int cloudsc_field_owner(size_t elem_size, int rank, const int64_t *lbounds, const int64_t *ubounds, int delayed,
                        const void *init_value, cloudsc_field **out);
/* Device / host data with the copy-on-stale rule above; stream = the field API's QUEUE (NULL: default stream,
 * synchronous).  The returned pointer addresses element (lbounds) of the whole array. */
int cloudsc_field_get_device_data(cloudsc_field *f, int mode, void **dev, void *stream);
int cloudsc_field_get_host_data(cloudsc_field *f, int mode, void **host, void *stream);
int cloudsc_field_sync_device(cloudsc_field *f, int mode, void *stream);
int cloudsc_field_sync_host(cloudsc_field *f, int mode, void *stream);
/* Host slice of last-dimension index `block` (lbound..ubound of the last dimension). */
int cloudsc_field_get_view(cloudsc_field *f, int64_t block, void **view);
/* Free the device copy (DELETE_DEVICE); the host copy keeps its status. */
int cloudsc_field_delete_device(cloudsc_field *f);
int cloudsc_field_status(const cloudsc_field *f, int *status);
/* GPU_STATS: number of host->device and device->host transfers so far. */
int cloudsc_field_stats(const cloudsc_field *f, int64_t *h2d, int64_t *d2h);
/* FIELD_DELETE: wrappers first sync the host copy; owners free their host memory.  NULL is accepted. */
int cloudsc_field_delete(cloudsc_field *f);

/* The cloud scheme of one physics step on fields: CLOUDSC's blocked arrays (the arrays of cloudsc_blocked_*_c,
 * in that order) and, for the fused CALLPAR sequence, the CLOUD_SATADJ inputs and TENDENCY_SATADJ outputs.
 * Field shapes are the blocked layout, Fortran (NPROMA, KLEV[+1], [n,] NGPBLKS). */
typedef struct cloudsc_step_fields {
  cloudsc_field *pt, *pq, *ptendency_cml_t, *ptendency_cml_q, *ptendency_cml_a, *ptendency_cml_cld;
  cloudsc_field *ptendency_loc_t, *ptendency_loc_q, *ptendency_loc_a, *ptendency_loc_cld;
  cloudsc_field *pvfl, *pvfi, *pap, *paph, *plsm, *ldcum, *ktype, *peis, *plu, *plude, *pludeli, *psnde;
  cloudsc_field *pmfu, *pmfd, *pa, *pclv, *pcovptot, *prainfrac_toprfz;
  cloudsc_field *pfsqlf, *pfsqif, *pfcqnng, *pfcqlng, *pfsqrf, *pfsqsf, *pfcqrng, *pfcqsng, *pfsqltur, *pfsqitur;
  cloudsc_field *pfplsl, *pfplsn, *pfhpsl, *pfhpsn;
  /* CLOUD_SATADJ (all NULL: CLOUDSC alone) */
  cloudsc_field *phrsw, *phrlw, *ptendency_vdf_t, *ptendency_vdf_q, *pvervel;
  /* TENDENCY_SATADJ (all NULL: not stored) */
  cloudsc_field *ptendency_satadj_t, *ptendency_satadj_q, *ptendency_satadj_a, *ptendency_satadj_cld;
} cloudsc_step_fields;

/* One launch over all blocks, replacing the per-block CLOUD_SATADJ / CLOUD_LAYER calls inside EC_PHYS_DRV's
 * OpenMP block loop (ec_phys_drv.F90:430-441; callpar.F90:1253-1296).  Inputs are taken with
 * get_device_data(RD), outputs and the in-place CML tendencies with get_device_data(RDWR), so host copies are
 * sent only when stale and results stay on the device until a get_host_data / sync_host.  precision: 0 FP64,
 * 1 FP32, 2 mixed (field element size must match).  Asynchronous on `stream`. */
int cloudsc_step_fields_run(const cloudsc_params_t *ydparams, int precision, int ngptot, double ptsphy,
                            const cloudsc_step_fields *fields, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CLOUDSC_FIELD_H */
