/*
 * cloudsc_tune.h -- launch configuration of the blocked CLOUDSC entries, set by the project's autotuner
 * (openifs_amd/autotune.py).  No counterpart in the reference: the IFS chooses only NPROMA.
 *
 * variant: 0 = cloudsc_blocked_dp_c, 1 = cloudsc_blocked_sp_c, 2 = cloudsc_blocked_mixed_c.
 * The work-group size (64, 128, 256 or 512 columns) applies to launches with the given KLEV and the default
 * configuration constants; 256 is used where nothing is set.  Process-wide; safe to call from any thread.
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

#ifdef __cplusplus
}
#endif

#endif /* CLOUDSC_TUNE_H */
