// SAR range-Doppler processor object.
#pragma once
#include "runtime.hpp"

struct mifft_rda {
    mifft_rda_params p{};
    mifft_plan* range_plan = nullptr;  // length n_range (twiddles, chirp FFT)
    mifft_plan* az_plan = nullptr;     // length n_azimuth
    mifft::DeviceBuffer range_filter;  // conj(FFT(chirp)), complex64
    mifft::DeviceBuffer az_a, az_d;    // FP64 azimuth phase model per Doppler bin
    mifft::DeviceBuffer az_table;      // azimuth filter factor tables E1 (Nr/64 rows), E2 (64 rows), FP32
    double az_b = 0.0;                 // cycles per range bin per unit D
    double range_scale = 1.0, az_scale = 1.0;
    int64_t chirp_samples = 0;
    mifft::DeviceBuffer rcmc_si, rcmc_w;  // RCMC: integer shift and 8 weights per Doppler bin
    mifft::DeviceBuffer work;          // one frame: range-compressed data (process_device), the
                                       // range-Doppler frame before RCMC
    hipStream_t stream = nullptr;
};
