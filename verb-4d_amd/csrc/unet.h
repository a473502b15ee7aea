// The learned warm start: a 4-level U-Net on the (E, alpha) grid of one
// (P,R) node, BF16 operands on MFMA with FP32 accumulation (unet.hip).
#pragma once
#include <cstddef>
#include <cstdint>

namespace verb {
namespace unet {

constexpr int kInCh = 16;    // 5 feature channels, zero-padded to one 16-wide MFMA k-chunk
constexpr int kConvs = 8;    // 3x3 convolutions
constexpr int kHeadCh = 32;  // channels into the 1x1 output convolution

// One 3x3 convolution: its input is channels [0, c0) from src0 (read at half
// resolution and nearest-upsampled when up0) followed by channels [c0,
// c0+c1) from src1; `level` is the resolution of its output (0 = full,
// each level halves height and width).
struct ConvSpec {
    int c0, up0, c1, cout, level;
};

// e1a, e1b | pool | e2 | pool | e3 | pool | bottleneck | up+skip d3, d2, d1
constexpr ConvSpec kSpec[kConvs] = {
    {kInCh, 0, 0, 32, 0}, {32, 0, 0, 32, 0}, {32, 0, 0, 64, 1},  {64, 0, 0, 96, 2},
    {96, 0, 0, 128, 3},   {128, 1, 96, 96, 2}, {96, 1, 64, 64, 1}, {64, 1, 32, 32, 0},
};

// FP32 parameter vector: for each convolution W [9][cin][cout] (tap-major,
// tap = 3*(dh+1) + (dw+1)) then b [cout]; then the head w [32], b [1].
inline size_t param_offset(int layer) {
    size_t o = 0;
    for (int l = 0; l < layer; l++) {
        const int cin = kSpec[l].c0 + kSpec[l].c1;
        o += (size_t)9 * cin * kSpec[l].cout + kSpec[l].cout;
    }
    return o;
}
inline size_t param_count() { return param_offset(kConvs) + kHeadCh + 1; }

}  // namespace unet
}  // namespace verb
