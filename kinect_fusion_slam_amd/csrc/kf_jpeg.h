// kf_jpeg.h - host side of the colour decode (baseline JPEG).
//
// The host parses the markers and runs the Huffman decoding, the one serial
// part of a JPEG decode, into coefficient blocks; dequantisation, the
// inverse DCT, chroma upsampling and colour conversion run on the device
// (kf_color.hip).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <string>

#define KF_JPEG_MAX_COMPONENTS 4

struct kf_jpeg_component {
  int h = 1, v = 1;      // sampling factors
  int tq = 0;            // quantisation table
  int dw = 0, dh = 0;    // downsampled size in samples (the real samples)
  int bw = 0, bh = 0;    // blocks stored per row / column (whole MCUs)
  int64_t block0 = 0;    // first block of the component in the coefficient array
};

struct kf_jpeg_frame {
  int H = 0, W = 0;
  int nc = 0;            // 1 (grey) or 3
  int hmax = 1, vmax = 1;
  bool rgb = false;      // 3 components that are R, G, B (no colour transform)
  kf_jpeg_component comp[KF_JPEG_MAX_COMPONENTS];
  uint16_t q[4][64] = {};// quantisation tables, natural order
  int64_t n_blocks = 0;  // blocks of all components
};

// Parses a baseline (SOF0 / SOF1, 8-bit, Huffman) JPEG.  With coef == NULL
// only the frame header is read (sizes, components, block counts).  With
// coef (capacity >= f->n_blocks blocks of 64 int16 each) every scan is
// decoded into it: block b of component c at coef[(comp[c].block0 + b) *
// 64], coefficients in natural order, not dequantised; blocks no scan
// covers stay zero.  Returns 0, or -1 with a message in *err.
int kf_jpeg_decode_entropy(const uint8_t* data, size_t n, kf_jpeg_frame* f, int16_t* coef,
                           int64_t capacity, std::string* err);
