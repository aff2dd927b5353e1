// kf_oracle_jpeg.cpp - CPU oracle for the colour path.
//
// TEST INFRASTRUCTURE ONLY.  Loaded by tests/ as the checker; the product
// library never links, loads or calls it.
//
// The reference loads each colour frame with cv::imread(path) (IMREAD_COLOR,
// system_utility.cpp:70-75), i.e. libjpeg-turbo's decompressor with its
// default parameters behind OpenCV's JPEG codec; neither is vendored in the
// reference.  This file restates that decoder's published algorithm for
// baseline sequential Huffman JPEG (ITU-T T.81 / JFIF 1.02 with the IJG
// conventions libjpeg-turbo keeps):
//   * entropy decoding: the DECODE / RECEIVE / EXTEND procedures of T.81
//     F.2.2, one bit at a time, restart markers (DRI) honoured; past the end
//     of the entropy-coded data the bit reader supplies zeros;
//   * dequantisation and the IJG "ISLOW" inverse DCT (jidctint.c's
//     Loeffler-Ligtenberg-Moschytz factorisation: 13-bit constants, 2 extra
//     bits kept between the column and row passes, outputs clamped by the
//     range-limit table that maps x to clamp(x + 128, 0, 255) for x in
//     [-512, 511] after masking x with 1023);
//   * "fancy" chroma upsampling (jdsample.c): for 2:1 horizontal factors the
//     3/4-1/4 triangle filter with rounding biases 1 and 2, for 2:1 in both
//     directions the separable triangle filter over column sums with biases
//     8 and 7, rows above the first and below the last real row replicated;
//     a component whose downsampled width is 2 or less is box-upsampled;
//     other integral factors replicate samples, except 1:2 vertical-only
//     (4:4:0), which is rejected;
//   * YCbCr -> RGB with jdcolor.c's 16-bit fixed-point tables
//     (R = Y + 1.40200 Cr, G = Y - 0.34414 Cb - 0.71414 Cr, B = Y + 1.77200 Cb);
//     without a JFIF marker, a 3-component image with an Adobe marker of
//     transform 0, or with no Adobe marker and component ids 'R', 'G', 'B',
//     is taken as RGB; one component is grey, replicated into B, G and R;
//   * output in OpenCV's channel order, B, G, R, H x W x 3 bytes.
// Progressive, arithmetic-coded, lossless and 12-bit JPEGs are rejected.
//
// The reference's point cloud (ThreeDViewer::SetUpPointClouds,
// threeD_viewer.cpp:20-55) is restated by oracle_point_cloud.
#include <stdint.h>
#include <string.h>

#include <vector>

namespace {

const int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                         12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                         35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                         58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// T.81 Annex C / F.2.2.3: MINCODE, MAXCODE, VALPTR per code length.
struct Huff {
  bool defined = false;
  int mincode[17], maxcode[18], valptr[17];
  uint8_t vals[256];
};

bool build_huff(const uint8_t bits[16], const uint8_t* vals, int nvals, Huff& h) {
  int total = 0;
  for (int i = 0; i < 16; ++i) total += bits[i];
  if (total > 256 || total != nvals) return false;
  memcpy(h.vals, vals, nvals);
  int code = 0, k = 0;
  for (int l = 1; l <= 16; ++l) {
    const int n = bits[l - 1];
    if (n) {
      h.valptr[l] = k;
      h.mincode[l] = code;
      code += n;
      k += n;
      h.maxcode[l] = code - 1;
    } else {
      h.maxcode[l] = -1;
    }
    code <<= 1;
  }
  h.maxcode[17] = 0x7fffffff;
  h.defined = true;
  return true;
}

struct Component {
  int id = 0, h = 1, v = 1, tq = 0;
  int td = 0, ta = 0;              // Huffman tables of the current scan
  int bw = 0, bh = 0;              // blocks per row / column stored
  int dw = 0, dh = 0;              // downsampled width / height in samples
  std::vector<int16_t> coef;       // bw * bh * 64, natural order, not dequantised
  int pred = 0;
};

struct Reader {
  const uint8_t* p;
  int64_t n, pos;
  uint32_t acc = 0;
  int nbits = 0;
  bool marker_hit = false;  // a marker ended the entropy-coded data
  int bit() {
    if (nbits == 0) {
      int byte = 0;
      if (!marker_hit && pos < n) {
        byte = p[pos];
        if (byte == 0xFF) {
          const int next = pos + 1 < n ? p[pos + 1] : -1;
          if (next == 0x00) {
            pos += 2;
          } else {
            marker_hit = true;  // leave the marker for the caller
            byte = 0;
          }
        } else {
          pos += 1;
        }
      }
      acc = (uint32_t)byte;
      nbits = 8;
    }
    --nbits;
    return (acc >> nbits) & 1;
  }
  int receive(int s) {
    int v = 0;
    for (int i = 0; i < s; ++i) v = (v << 1) | bit();
    return v;
  }
  // T.81 F.2.2.3 DECODE
  int decode(const Huff& h) {
    int code = bit();
    int l = 1;
    while (code > h.maxcode[l]) {
      code = (code << 1) | bit();
      if (++l > 16) return -1;
    }
    return h.vals[h.valptr[l] + code - h.mincode[l]];
  }
  // byte alignment and the RSTn marker at a restart boundary
  bool restart() {
    nbits = 0;
    marker_hit = false;
    while (pos + 1 < n && !(p[pos] == 0xFF && p[pos + 1] >= 0xD0 && p[pos + 1] <= 0xD7)) ++pos;
    if (pos + 1 >= n) return false;
    pos += 2;
    return true;
  }
};

inline int extend(int v, int s) { return s == 0 ? 0 : (v < (1 << (s - 1)) ? v - (1 << s) + 1 : v); }

// jidctint.c's constants: round(c * 2^13)
const int32_t F0_298631336 = 2446, F0_390180644 = 3196, F0_541196100 = 4433,
              F0_765366865 = 6270, F0_899976223 = 7373, F1_175875602 = 9633,
              F1_501321110 = 12299, F1_847759065 = 15137, F1_961570560 = 16069,
              F2_053119869 = 16819, F2_562915447 = 20995, F3_072711026 = 25172;

inline int32_t descale(int64_t x, int n) { return (int32_t)((x + ((int64_t)1 << (n - 1))) >> n); }

// the post-IDCT range limit: x masked to 10 bits, then clamp(x + 128) for
// x in [-512, 511]
inline uint8_t range_limit(int32_t x) {
  int32_t m = x & 1023;
  if (m >= 512) m -= 1024;
  const int32_t v = m + 128;
  return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
}

// One 8x8 block: coef (natural order) times q, ISLOW IDCT, 8 rows of 8
// samples at out with the given stride.
void idct_islow(const int16_t* coef, const uint16_t* q, uint8_t* out, int stride) {
  const int CB = 13, P1 = 2;
  int32_t ws[64];
  for (int c = 0; c < 8; ++c) {
    int32_t in[8];
    for (int r = 0; r < 8; ++r) in[r] = (int32_t)coef[r * 8 + c] * (int32_t)q[r * 8 + c];
    if (!in[1] && !in[2] && !in[3] && !in[4] && !in[5] && !in[6] && !in[7]) {
      const int32_t dc = in[0] * (1 << P1);
      for (int r = 0; r < 8; ++r) ws[r * 8 + c] = dc;
      continue;
    }
    // even part
    int64_t z2 = in[2], z3 = in[6];
    int64_t z1 = (z2 + z3) * F0_541196100;
    const int64_t t2 = z1 + z3 * -F1_847759065;
    const int64_t t3 = z1 + z2 * F0_765366865;
    z2 = in[0];
    z3 = in[4];
    const int64_t t0 = (z2 + z3) * (1 << CB);
    const int64_t t1 = (z2 - z3) * (1 << CB);
    const int64_t e10 = t0 + t3, e13 = t0 - t3, e11 = t1 + t2, e12 = t1 - t2;
    // odd part
    int64_t o0 = in[7], o1 = in[5], o2 = in[3], o3 = in[1];
    z1 = o0 + o3;
    z2 = o1 + o2;
    z3 = o0 + o2;
    int64_t z4 = o1 + o3;
    const int64_t z5 = (z3 + z4) * F1_175875602;
    o0 *= F0_298631336;
    o1 *= F2_053119869;
    o2 *= F3_072711026;
    o3 *= F1_501321110;
    z1 *= -F0_899976223;
    z2 *= -F2_562915447;
    z3 *= -F1_961570560;
    z4 *= -F0_390180644;
    z3 += z5;
    z4 += z5;
    o0 += z1 + z3;
    o1 += z2 + z4;
    o2 += z2 + z3;
    o3 += z1 + z4;
    ws[0 * 8 + c] = descale(e10 + o3, CB - P1);
    ws[7 * 8 + c] = descale(e10 - o3, CB - P1);
    ws[1 * 8 + c] = descale(e11 + o2, CB - P1);
    ws[6 * 8 + c] = descale(e11 - o2, CB - P1);
    ws[2 * 8 + c] = descale(e12 + o1, CB - P1);
    ws[5 * 8 + c] = descale(e12 - o1, CB - P1);
    ws[3 * 8 + c] = descale(e13 + o0, CB - P1);
    ws[4 * 8 + c] = descale(e13 - o0, CB - P1);
  }
  for (int r = 0; r < 8; ++r) {
    const int32_t* w = ws + r * 8;
    uint8_t* o = out + (size_t)r * stride;
    if (!w[1] && !w[2] && !w[3] && !w[4] && !w[5] && !w[6] && !w[7]) {
      const uint8_t v = range_limit(descale(w[0], P1 + 3));
      for (int k = 0; k < 8; ++k) o[k] = v;
      continue;
    }
    int64_t z2 = w[2], z3 = w[6];
    int64_t z1 = (z2 + z3) * F0_541196100;
    const int64_t t2 = z1 + z3 * -F1_847759065;
    const int64_t t3 = z1 + z2 * F0_765366865;
    const int64_t t0 = ((int64_t)w[0] + w[4]) * (1 << CB);
    const int64_t t1 = ((int64_t)w[0] - w[4]) * (1 << CB);
    const int64_t e10 = t0 + t3, e13 = t0 - t3, e11 = t1 + t2, e12 = t1 - t2;
    int64_t o0 = w[7], o1 = w[5], o2 = w[3], o3 = w[1];
    z1 = o0 + o3;
    z2 = o1 + o2;
    z3 = o0 + o2;
    int64_t z4 = o1 + o3;
    const int64_t z5 = (z3 + z4) * F1_175875602;
    o0 *= F0_298631336;
    o1 *= F2_053119869;
    o2 *= F3_072711026;
    o3 *= F1_501321110;
    z1 *= -F0_899976223;
    z2 *= -F2_562915447;
    z3 *= -F1_961570560;
    z4 *= -F0_390180644;
    z3 += z5;
    z4 += z5;
    o0 += z1 + z3;
    o1 += z2 + z4;
    o2 += z2 + z3;
    o3 += z1 + z4;
    const int S = CB + P1 + 3;
    o[0] = range_limit(descale(e10 + o3, S));
    o[7] = range_limit(descale(e10 - o3, S));
    o[1] = range_limit(descale(e11 + o2, S));
    o[6] = range_limit(descale(e11 - o2, S));
    o[2] = range_limit(descale(e12 + o1, S));
    o[5] = range_limit(descale(e12 - o1, S));
    o[3] = range_limit(descale(e13 + o0, S));
    o[4] = range_limit(descale(e13 - o0, S));
  }
}

// Upsampling of one component plane (dw x dh real samples, row stride sw)
// to H x W by factors (fh, fv).
void upsample(const uint8_t* in, int sw, int dw, int dh, int fh, int fv, int H, int W,
              std::vector<uint8_t>& out) {
  out.assign((size_t)H * W, 0);
  auto at = [&](int x, int y) -> int { return in[(size_t)y * sw + x]; };
  const bool fancy_h = fh == 2 && dw > 2;
  if (fh == 2 && fv == 1 && fancy_h) {
    for (int y = 0; y < H; ++y) {
      std::vector<int> row(2 * dw);
      for (int i = 0; i < dw; ++i) {
        const int c3 = at(i, y) * 3;
        const int left = i == 0 ? -1 : at(i - 1, y), right = i == dw - 1 ? -1 : at(i + 1, y);
        row[2 * i] = i == 0 ? at(0, y) : (c3 + left + 1) >> 2;
        row[2 * i + 1] = i == dw - 1 ? at(i, y) : (c3 + right + 2) >> 2;
      }
      for (int x = 0; x < W; ++x) out[(size_t)y * W + x] = (uint8_t)row[x];
    }
    return;
  }
  if (fh == 2 && fv == 2 && fancy_h) {
    for (int oy = 0; oy < H; ++oy) {
      const int y = oy >> 1;
      // the nearer row is y, the further one is above (even output rows) or
      // below (odd), replicated at the first and last real rows
      const int y2 = (oy & 1) ? (y + 1 < dh ? y + 1 : dh - 1) : (y > 0 ? y - 1 : 0);
      std::vector<int> cs(dw);
      for (int i = 0; i < dw; ++i) cs[i] = at(i, y) * 3 + at(i, y2);
      std::vector<int> row(2 * dw);
      for (int i = 0; i < dw; ++i) {
        if (i == 0)
          row[0] = (cs[0] * 4 + 8) >> 4;
        else
          row[2 * i] = (cs[i] * 3 + cs[i - 1] + 8) >> 4;
        if (i == dw - 1)
          row[2 * i + 1] = (cs[i] * 4 + 7) >> 4;
        else
          row[2 * i + 1] = (cs[i] * 3 + cs[i + 1] + 7) >> 4;
      }
      for (int x = 0; x < W; ++x) out[(size_t)oy * W + x] = (uint8_t)row[x];
    }
    return;
  }
  // full size or box replication
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) out[(size_t)y * W + x] = (uint8_t)at(x / fh, y / fv);
}

inline uint8_t clamp255(int v) { return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

int16_t kSizeOnlyStorage;
int16_t* const kSizeOnly = &kSizeOnlyStorage;  // coefficient request: count only

}  // namespace

extern "C" {

// Decodes a baseline JPEG held in memory.  Returns 0 and the size in *H, *W
// (bgr NULL: size only), or -1 (malformed or unsupported: progressive,
// arithmetic, lossless, 12-bit, more than 4 components, sampling factors
// that are not integral divisors of the largest).
static int decode(const uint8_t* data, int64_t n, int* H, int* W, uint8_t* bgr,
                  int16_t* coef_out, int64_t* coef_out_blocks);

int oracle_jpeg_decode(const uint8_t* data, int64_t n, int* H, int* W, uint8_t* bgr) {
  return decode(data, n, H, W, bgr, nullptr, nullptr);
}

// The entropy-decoded coefficient blocks (natural order, not dequantised),
// component after component, each over its whole-MCU block grid in raster
// order.  coef NULL: the block count only.
int oracle_jpeg_coefficients(const uint8_t* data, int64_t n, int16_t* coef, int64_t* n_blocks) {
  int h = 0, w = 0;
  return decode(data, n, &h, &w, nullptr, coef ? coef : kSizeOnly, n_blocks);
}

}  // extern "C"

static int decode(const uint8_t* data, int64_t n, int* H, int* W, uint8_t* bgr,
                  int16_t* coef_out, int64_t* coef_out_blocks) {
  if (!data || n < 4 || data[0] != 0xFF || data[1] != 0xD8) return -1;
  uint16_t qt[4][64] = {};
  Huff hdc[4], hac[4];
  std::vector<Component> comps;
  int width = 0, height = 0, restart = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
  bool jfif = false, adobe = false, frame = false, any_scan = false;
  int adobe_transform = -1;
  int64_t pos = 2;
  while (pos + 4 <= n) {
    if (data[pos] != 0xFF) return -1;
    const int m = data[pos + 1];
    if (m == 0xFF) {
      ++pos;
      continue;
    }
    if (m == 0xD9) break;  // EOI
    const int len = (data[pos + 2] << 8) | data[pos + 3];
    if (len < 2 || pos + 2 + len > n) return -1;
    const uint8_t* s = data + pos + 4;
    const int sl = len - 2;
    if (m == 0xDB) {  // DQT
      int k = 0;
      while (k < sl) {
        const int pq = s[k] >> 4, tq = s[k] & 15;
        if (tq > 3 || pq > 1) return -1;
        ++k;
        for (int i = 0; i < 64; ++i) {
          const int v = pq ? (s[k] << 8) | s[k + 1] : s[k];
          k += pq ? 2 : 1;
          qt[tq][kZigzag[i]] = (uint16_t)v;
        }
      }
    } else if (m == 0xC4) {  // DHT
      int k = 0;
      while (k + 17 <= sl) {
        const int tc = s[k] >> 4, th = s[k] & 15;
        if (tc > 1 || th > 3) return -1;
        uint8_t bits[16];
        int total = 0;
        for (int i = 0; i < 16; ++i) total += bits[i] = s[k + 1 + i];
        if (k + 17 + total > sl) return -1;
        if (!build_huff(bits, s + k + 17, total, tc ? hac[th] : hdc[th])) return -1;
        k += 17 + total;
      }
    } else if (m == 0xC0 || m == 0xC1) {  // SOF0 / SOF1 (Huffman, sequential)
      if (sl < 6 || s[0] != 8) return -1;
      height = (s[1] << 8) | s[2];
      width = (s[3] << 8) | s[4];
      const int nf = s[5];
      if (nf < 1 || nf > 4 || sl < 6 + 3 * nf || width < 1 || height < 1) return -1;
      comps.resize(nf);
      for (int i = 0; i < nf; ++i) {
        comps[i].id = s[6 + 3 * i];
        comps[i].h = s[7 + 3 * i] >> 4;
        comps[i].v = s[7 + 3 * i] & 15;
        comps[i].tq = s[8 + 3 * i];
        if (comps[i].h < 1 || comps[i].h > 4 || comps[i].v < 1 || comps[i].v > 4 || comps[i].tq > 3)
          return -1;
        hmax = comps[i].h > hmax ? comps[i].h : hmax;
        vmax = comps[i].v > vmax ? comps[i].v : vmax;
      }
      mcux = (width + 8 * hmax - 1) / (8 * hmax);
      mcuy = (height + 8 * vmax - 1) / (8 * vmax);
      for (Component& c : comps) {
        if (hmax % c.h || vmax % c.v) return -1;
        // 1:2 vertical-only chroma (4:4:0) has a filter of its own in
        // libjpeg-turbo that no test here can pin: rejected
        if (hmax / c.h == 1 && vmax / c.v == 2) return -1;
        c.dw = (width * c.h + hmax - 1) / hmax;
        c.dh = (height * c.v + vmax - 1) / vmax;
        c.bw = mcux * c.h;
        c.bh = mcuy * c.v;
        c.coef.assign((size_t)c.bw * c.bh * 64, 0);
      }
      frame = true;
      if (H) *H = height;
      if (W) *W = width;
      if (!bgr && !coef_out) return 0;
    } else if ((m >= 0xC2 && m <= 0xC3) || (m >= 0xC5 && m <= 0xCF && m != 0xC8 && m != 0xCC)) {
      return -1;  // progressive, lossless, hierarchical or arithmetic
    } else if (m == 0xDD) {  // DRI
      if (sl < 2) return -1;
      restart = (s[0] << 8) | s[1];
    } else if (m == 0xE0) {
      if (sl >= 5 && !memcmp(s, "JFIF\0", 5)) jfif = true;
    } else if (m == 0xEE) {
      if (sl >= 12 && !memcmp(s, "Adobe", 5)) {
        adobe = true;
        adobe_transform = s[11];
      }
    } else if (m == 0xDA) {  // SOS
      if (!frame) return -1;
      const int ns = s[0];
      if (ns < 1 || ns > 4 || sl < 1 + 2 * ns + 3) return -1;
      std::vector<Component*> sc;
      for (int i = 0; i < ns; ++i) {
        Component* c = nullptr;
        for (Component& cc : comps)
          if (cc.id == s[1 + 2 * i]) c = &cc;
        if (!c) return -1;
        c->td = s[2 + 2 * i] >> 4;
        c->ta = s[2 + 2 * i] & 15;
        if (c->td > 3 || c->ta > 3 || !hdc[c->td].defined || !hac[c->ta].defined) return -1;
        c->pred = 0;
        sc.push_back(c);
      }
      Reader rd{data, n, pos + 2 + len};
      // MCUs of this scan: interleaved (the frame's MCU grid) or one block
      // of the component at a time over its own block grid
      const bool inter = ns > 1;
      const int cols = inter ? mcux : (sc[0]->dw + 7) / 8;
      const int rows = inter ? mcuy : (sc[0]->dh + 7) / 8;
      int64_t left = restart;
      for (int my = 0; my < rows; ++my)
        for (int mx = 0; mx < cols; ++mx) {
          if (restart && left == 0) {
            if (!rd.restart()) return -1;
            for (Component* c : sc) c->pred = 0;
            left = restart;
          }
          for (Component* c : sc) {
            const int nby = inter ? c->v : 1, nbx = inter ? c->h : 1;
            for (int by = 0; by < nby; ++by)
              for (int bx = 0; bx < nbx; ++bx) {
                const int gx = inter ? mx * c->h + bx : mx, gy = inter ? my * c->v + by : my;
                int16_t* blk = &c->coef[((size_t)gy * c->bw + gx) * 64];
                const int t = rd.decode(hdc[c->td]);
                if (t < 0 || t > 11) return -1;
                c->pred += extend(rd.receive(t), t);
                blk[0] = (int16_t)c->pred;
                for (int k = 1; k < 64;) {
                  const int rs = rd.decode(hac[c->ta]);
                  if (rs < 0) return -1;
                  const int r = rs >> 4, sz = rs & 15;
                  if (sz) {
                    k += r;
                    if (k > 63) return -1;
                    blk[kZigzag[k]] = (int16_t)extend(rd.receive(sz), sz);
                    ++k;
                  } else if (r == 15) {
                    k += 16;
                  } else {
                    break;
                  }
                }
              }
          }
          if (restart) --left;
        }
      any_scan = true;
      // continue after the entropy-coded data: the next marker
      int64_t q = rd.pos;
      while (q + 1 < n && !(data[q] == 0xFF && data[q + 1] != 0x00 &&
                            !(data[q + 1] >= 0xD0 && data[q + 1] <= 0xD7)))
        ++q;
      pos = q;
      continue;
    }
    pos += 2 + len;
  }
  if (coef_out) {
    // the coefficient blocks, component after component
    int64_t total = 0;
    for (const Component& c : comps) total += (int64_t)c.bw * c.bh;
    *coef_out_blocks = total;
    if (coef_out != kSizeOnly) {
      int64_t k = 0;
      for (const Component& c : comps) {
        memcpy(coef_out + k * 64, c.coef.data(), c.coef.size() * sizeof(int16_t));
        k += (int64_t)c.bw * c.bh;
      }
    }
    return any_scan ? 0 : -1;
  }
  if (!frame || !any_scan || !bgr) return frame ? 0 : -1;
  // dequantise + IDCT, upsample
  const int nc = (int)comps.size();
  std::vector<std::vector<uint8_t>> full(nc);
  for (int ci = 0; ci < nc; ++ci) {
    Component& c = comps[ci];
    const int sw = c.bw * 8;
    std::vector<uint8_t> plane((size_t)sw * c.bh * 8);
    for (int by = 0; by < c.bh; ++by)
      for (int bx = 0; bx < c.bw; ++bx)
        idct_islow(&c.coef[((size_t)by * c.bw + bx) * 64], qt[c.tq],
                   &plane[(size_t)by * 8 * sw + bx * 8], sw);
    upsample(plane.data(), sw, c.dw, c.dh, hmax / c.h, vmax / c.v, height, width, full[ci]);
  }
  // colour conversion
  bool rgb = false;
  if (nc == 3) {
    // a JFIF marker first, then an Adobe marker's transform, then the ids
    if (jfif)
      rgb = false;
    else if (adobe)
      rgb = adobe_transform == 0;
    else
      rgb = comps[0].id == 'R' && comps[1].id == 'G' && comps[2].id == 'B';
  } else if (nc != 1) {
    return -1;  // CMYK / YCCK are not handled
  }
  const int64_t one_half = 1 << 15;
  auto fix = [](double v) { return (int64_t)(v * 65536.0 + 0.5); };
  for (int64_t i = 0; i < (int64_t)height * width; ++i) {
    uint8_t* o = bgr + 3 * i;
    if (nc == 1) {
      o[0] = o[1] = o[2] = full[0][i];
    } else if (rgb) {
      o[2] = full[0][i];
      o[1] = full[1][i];
      o[0] = full[2][i];
    } else {
      const int y = full[0][i];
      const int64_t cb = (int64_t)full[1][i] - 128, cr = (int64_t)full[2][i] - 128;
      const int r = y + (int)((fix(1.40200) * cr + one_half) >> 16);
      const int b = y + (int)((fix(1.77200) * cb + one_half) >> 16);
      const int g = y + (int)((-fix(0.34414) * cb + -fix(0.71414) * cr + one_half) >> 16);
      o[2] = clamp255(r);
      o[1] = clamp255(g);
      o[0] = clamp255(b);
    }
  }
  return 0;
}

extern "C" {

// ThreeDViewer::SetUpPointClouds (threeD_viewer.cpp:20-55): for pixel
// index = u * W + v, where the raycast vertex z != 0, the point (x, y, z)
// and the colour of the same pixel of the BGR image packed as
// r << 16 | g << 8 | b and stored as the bits of a float; elsewhere zeros.
// out: H*W records of 4 floats {x, y, z, rgb}.
void oracle_point_cloud(const float* x, const float* y, const float* z, const uint8_t* bgr,
                        int H, int W, float* out) {
  for (int64_t i = 0; i < (int64_t)H * W; ++i) {
    float* o = out + 4 * i;
    if (z[i] != 0.f) {
      const uint32_t rgb = (uint32_t)bgr[3 * i + 2] << 16 | (uint32_t)bgr[3 * i + 1] << 8 |
                           (uint32_t)bgr[3 * i];
      float f;
      memcpy(&f, &rgb, 4);
      o[0] = x[i];
      o[1] = y[i];
      o[2] = z[i];
      o[3] = f;
    } else {
      const uint32_t zero = 0;
      float f;
      memcpy(&f, &zero, 4);
      o[0] = o[1] = o[2] = 0.f;
      o[3] = f;
    }
  }
}

}  // extern "C"
