/* Prints sizeof / offsetof of the C ABI structs of include/lsim.h as "name value" lines, so the
 * host test can compare them with the ctypes declarations of larndsim_amd/_lib.py. */
#include <stddef.h>
#include <stdio.h>

#include "lsim.h"

#define SZ(T) printf("sizeof.%s %zu\n", #T, sizeof(T))
#define OFF(T, f) printf("offsetof.%s.%s %zu\n", #T, #f, offsetof(T, f))

int main(void) {
    SZ(lsim_geometry);
    OFF(lsim_geometry, n_tpc); OFF(lsim_geometry, n_pixels_x); OFF(lsim_geometry, n_pixels_y);
    OFF(lsim_geometry, pixel_pitch); OFF(lsim_geometry, tpc_borders);
    SZ(lsim_config);
    OFF(lsim_config, number_pix_neighbors); OFF(lsim_config, nb_tran_diff_bins);
    OFF(lsim_config, nb_sampling_bins_per_pixel); OFF(lsim_config, signal_length); OFF(lsim_config, n_ticks);
    OFF(lsim_config, max_adc_values); OFF(lsim_config, adc_interval); OFF(lsim_config, n_templates);
    OFF(lsim_config, long_diff_extent); OFF(lsim_config, long_diff_template);
    SZ(lsim_params);
    OFF(lsim_params, eField); OFF(lsim_params, Ab); OFF(lsim_params, kb); OFF(lsim_params, lifetime);
    OFF(lsim_params, vdrift); OFF(lsim_params, long_diff); OFF(lsim_params, tran_diff);
    OFF(lsim_params, lArDensity); OFF(lsim_params, alpha); OFF(lsim_params, beta); OFF(lsim_params, R_param);
    OFF(lsim_params, MeVToElectrons); OFF(lsim_params, vdrift_eff); OFF(lsim_params, dvdrift_deField);
    OFF(lsim_params, shift_x); OFF(lsim_params, shift_y); OFF(lsim_params, shift_z); OFF(lsim_params, size_margin);
    OFF(lsim_params, t_sampling); OFF(lsim_params, threshold); OFF(lsim_params, gain); OFF(lsim_params, v_pedestal);
    OFF(lsim_params, v_cm); OFF(lsim_params, v_ref); OFF(lsim_params, adc_counts);
    OFF(lsim_params, recombination_mode); OFF(lsim_params, pad_);
    SZ(lsim_segments);
    OFF(lsim_segments, x); OFF(lsim_segments, y); OFF(lsim_segments, z); OFF(lsim_segments, dE);
    OFF(lsim_segments, dEdx); OFF(lsim_segments, event_id); OFF(lsim_segments, z_start); OFF(lsim_segments, z_end);
    OFF(lsim_segments, dx); OFF(lsim_segments, stride); OFF(lsim_segments, n);
    SZ(lsim_fwd_out);
    OFF(lsim_fwd_out, n_rows); OFF(lsim_fwd_out, pixel_ids); OFF(lsim_fwd_out, wfs); OFF(lsim_fwd_out, adc);
    OFF(lsim_fwd_out, ticks); OFF(lsim_fwd_out, tape); OFF(lsim_fwd_out, n_hits); OFF(lsim_fwd_out, hit_adc);
    OFF(lsim_fwd_out, hit_x); OFF(lsim_fwd_out, hit_y); OFF(lsim_fwd_out, hit_z); OFF(lsim_fwd_out, hit_tick);
    OFF(lsim_fwd_out, hit_prob); OFF(lsim_fwd_out, hit_event); OFF(lsim_fwd_out, hit_pixel);
    OFF(lsim_fwd_out, hit_index); OFF(lsim_fwd_out, status);
    SZ(lsim_chop_layout);
    OFF(lsim_chop_layout, n_fields); OFF(lsim_chop_layout, start); OFF(lsim_chop_layout, end);
    OFF(lsim_chop_layout, mid); OFF(lsim_chop_layout, dE); OFF(lsim_chop_layout, dx); OFF(lsim_chop_layout, event);
    OFF(lsim_chop_layout, n_zero); OFF(lsim_chop_layout, zero); OFF(lsim_chop_layout, n_neg_one);
    OFF(lsim_chop_layout, neg_one);
    return 0;
}
