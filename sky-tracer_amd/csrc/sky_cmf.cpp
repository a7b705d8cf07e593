// sky_cmf.cpp -- CIE 1931 2-degree colour-matching functions, the table
// lookup of crates/sky-core/src/spectrum.rs:196-219 (cie_1931_2deg): clamped
// outside the table, exact at table wavelengths, linear in between, in FP32.
// The table is the CIE 1931 standard observer at 5 nm steps, 360-830 nm
// (published CIE data; the reference reads the same observer at 1 nm from
// data/CIE_xyz_1931_2deg.csv, so values agree at every multiple of 5 nm).
#include "../../include/sky_scan.h"
#include "sky_internal.hpp"

namespace {

constexpr int kCieCount = 95;
constexpr float kCieFirst = 360.0f, kCieStep = 5.0f;
constexpr float kCie[kCieCount][3] = {
    {0.0001299f, 3.917e-06f, 0.0006061f},
    {0.0002321f, 6.965e-06f, 0.001086f},
    {0.0004149f, 1.239e-05f, 0.001946f},
    {0.0007416f, 2.202e-05f, 0.003486f},
    {0.001368f, 3.9e-05f, 0.006450001f},
    {0.002236f, 6.4e-05f, 0.01054999f},
    {0.004243f, 0.00012f, 0.02005001f},
    {0.00765f, 0.000217f, 0.03621f},
    {0.01431f, 0.000396f, 0.06785001f},
    {0.02319f, 0.00064f, 0.1102f},
    {0.04351f, 0.00121f, 0.2074f},
    {0.07763f, 0.00218f, 0.3713f},
    {0.13438f, 0.004f, 0.6456f},
    {0.21477f, 0.0073f, 1.0390501f},
    {0.2839f, 0.0116f, 1.3856f},
    {0.3285f, 0.01684f, 1.62296f},
    {0.34828f, 0.023f, 1.74706f},
    {0.34806f, 0.0298f, 1.7826f},
    {0.3362f, 0.038f, 1.77211f},
    {0.3187f, 0.048f, 1.7441f},
    {0.2908f, 0.06f, 1.6692f},
    {0.2511f, 0.0739f, 1.5281f},
    {0.19536f, 0.09098f, 1.28764f},
    {0.1421f, 0.1126f, 1.0419f},
    {0.09564f, 0.13902f, 0.8129501f},
    {0.05795001f, 0.1693f, 0.6162f},
    {0.03201f, 0.20802f, 0.46518f},
    {0.0147f, 0.2586f, 0.3533f},
    {0.0049f, 0.323f, 0.272f},
    {0.0024f, 0.4073f, 0.2123f},
    {0.0093f, 0.503f, 0.1582f},
    {0.0291f, 0.6082f, 0.1117f},
    {0.06327f, 0.71f, 0.07824999f},
    {0.1096f, 0.7932f, 0.05725001f},
    {0.1655f, 0.862f, 0.04216f},
    {0.2257499f, 0.9148501f, 0.02984f},
    {0.2904f, 0.954f, 0.0203f},
    {0.3597f, 0.9803f, 0.0134f},
    {0.4334499f, 0.9949501f, 0.008749999f},
    {0.5120501f, 1.0f, 0.005749999f},
    {0.5945f, 0.995f, 0.0039f},
    {0.6784f, 0.9786f, 0.002749999f},
    {0.7621f, 0.952f, 0.0021f},
    {0.8425f, 0.9154f, 0.0018f},
    {0.9163f, 0.87f, 0.001650001f},
    {0.9786f, 0.8163f, 0.0014f},
    {1.0263f, 0.757f, 0.0011f},
    {1.0567f, 0.6949f, 0.001f},
    {1.0622f, 0.631f, 0.0008f},
    {1.0456f, 0.5668f, 0.0006f},
    {1.0026f, 0.503f, 0.00034f},
    {0.9384f, 0.4412f, 0.00024f},
    {0.8544499f, 0.381f, 0.00019f},
    {0.7514f, 0.321f, 0.0001f},
    {0.6424f, 0.265f, 4.999999e-05f},
    {0.5419f, 0.217f, 3e-05f},
    {0.4479f, 0.175f, 2e-05f},
    {0.3608f, 0.1382f, 1e-05f},
    {0.2835f, 0.107f, 0.0f},
    {0.2187f, 0.0816f, 0.0f},
    {0.1649f, 0.061f, 0.0f},
    {0.1212f, 0.04458f, 0.0f},
    {0.0874f, 0.032f, 0.0f},
    {0.0636f, 0.0232f, 0.0f},
    {0.04677f, 0.017f, 0.0f},
    {0.0329f, 0.01192f, 0.0f},
    {0.0227f, 0.00821f, 0.0f},
    {0.01584f, 0.005723f, 0.0f},
    {0.01135916f, 0.004102f, 0.0f},
    {0.008110916f, 0.002929f, 0.0f},
    {0.005790346f, 0.002091f, 0.0f},
    {0.004109457f, 0.001484f, 0.0f},
    {0.002899327f, 0.001047f, 0.0f},
    {0.00204919f, 0.00074f, 0.0f},
    {0.001439971f, 0.00052f, 0.0f},
    {0.0009999493f, 0.0003611f, 0.0f},
    {0.0006900786f, 0.0002492f, 0.0f},
    {0.0004760213f, 0.0001719f, 0.0f},
    {0.0003323011f, 0.00012f, 0.0f},
    {0.0002348261f, 8.48e-05f, 0.0f},
    {0.0001661505f, 6e-05f, 0.0f},
    {0.000117413f, 4.24e-05f, 0.0f},
    {8.307527e-05f, 3e-05f, 0.0f},
    {5.870652e-05f, 2.12e-05f, 0.0f},
    {4.150994e-05f, 1.499e-05f, 0.0f},
    {2.935326e-05f, 1.06e-05f, 0.0f},
    {2.067383e-05f, 7.4657e-06f, 0.0f},
    {1.455977e-05f, 5.2578e-06f, 0.0f},
    {1.025398e-05f, 3.7029e-06f, 0.0f},
    {7.221456e-06f, 2.6078e-06f, 0.0f},
    {5.085868e-06f, 1.8366e-06f, 0.0f},
    {3.581652e-06f, 1.2934e-06f, 0.0f},
    {2.522525e-06f, 9.1093e-07f, 0.0f},
    {1.776509e-06f, 6.4153e-07f, 0.0f},
    {1.251141e-06f, 4.5181e-07f, 0.0f},
};

}  // namespace

extern "C" sky_status sky_cie_1931_2deg(const float *lambda_nm, uint32_t n, double *xyz)
{
    if (n && (!lambda_nm || !xyz)) return sky::set_error(SKY_ERR_INVALID_CONFIG, "null argument");
    for (uint32_t i = 0; i < n; ++i) {
        const float l = lambda_nm[i];
        float v[3];
        if (!(l > kCieFirst)) {
            for (int c = 0; c < 3; ++c) v[c] = kCie[0][c];
        } else if (l >= kCieFirst + kCieStep * (kCieCount - 1)) {
            for (int c = 0; c < 3; ++c) v[c] = kCie[kCieCount - 1][c];
        } else {
            int k = (int)((l - kCieFirst) / kCieStep);
            const float a = kCieFirst + kCieStep * (float)k;
            if (l == a) {
                for (int c = 0; c < 3; ++c) v[c] = kCie[k][c];
            } else {
                const float t = (l - a) / kCieStep;
                for (int c = 0; c < 3; ++c) v[c] = kCie[k][c] + (kCie[k + 1][c] - kCie[k][c]) * t;
            }
        }
        for (int c = 0; c < 3; ++c) xyz[3 * i + c] = (double)v[c];
    }
    return SKY_OK;
}
