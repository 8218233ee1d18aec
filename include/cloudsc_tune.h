/*
 * cloudsc_tune.h -- launch configuration of the blocked CLOUDSC entries, set by the project's autotuner
 * (openifs_amd/autotune.py).  No counterpart in the reference: the IFS chooses only NPROMA.
 *
 * variant: 0 = cloudsc_blocked_dp_c, 1 = cloudsc_blocked_sp_c, 2 = cloudsc_blocked_mixed_c.
 * A launch configuration applies to launches with the given KLEV and the default configuration constants:
 *   block          columns per work-group, 64, 128, 256 or 512 (256 where nothing is set);
 *   waves_per_simd minimum waves per SIMD the kernel's register allocation must allow: 0 = the compiler's
 *                  choice (dp 1, sp 3, mixed 2 waves), or 2 (dp), 4 (sp), 3 or 4 (mixed), fewer registers per lane
 *                  and possibly spills;
 *   klev_compiled  1 = use the kernel with the level count compiled in (instantiated for KLEV 137 and 91;
 *                  other KLEV fall back to the run-time kernel).
 * waves_per_simd and klev_compiled are instantiated for blocks of 128 and 256.  Process-wide; safe to call from
 * any thread.
 */
#ifndef CLOUDSC_TUNE_H
#define CLOUDSC_TUNE_H

#ifdef __cplusplus
extern "C" {
#endif

/* block = 0 removes the setting.  Returns CLOUDSC_EDIMS for an unknown variant or block size. */
int cloudsc_tune_set_block(int variant, int klev, int block);
/* The work-group size launches of (variant, klev) use. */
int cloudsc_tune_get_block(int variant, int klev);
/* Full launch configuration; block = 0 removes the setting.  Returns CLOUDSC_EDIMS for a combination that is not
 * instantiated (see above). */
int cloudsc_tune_set_launch(int variant, int klev, int block, int waves_per_simd, int klev_compiled);
int cloudsc_tune_get_launch(int variant, int klev, int* block, int* waves_per_simd, int* klev_compiled);

#ifdef __cplusplus
}
#endif

#endif /* CLOUDSC_TUNE_H */
