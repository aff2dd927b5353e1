// kf_jpeg.cpp - marker parsing and Huffman decoding of baseline JPEG.
//
// The reference reads each colour frame with cv::imread (system_utility.cpp:
// 70-75), i.e. libjpeg-turbo behind OpenCV.  This file implements the
// entropy-decoding half of that decode for baseline sequential Huffman JPEG
// (ITU-T T.81): DQT, DHT, SOF0/SOF1, DRI, SOS, restart markers, interleaved
// and single-component scans, the JFIF / Adobe / component-id rule for
// whether 3 components are YCbCr or RGB.  Past the end of the entropy-coded
// data the bit reader supplies zeros, as libjpeg does.  Progressive,
// arithmetic-coded, lossless and 12-bit streams are rejected.
#include "kf_jpeg.h"

#include <string.h>

#include <vector>

namespace {

// natural-order index of the k-th coefficient in zig-zag order
const uint8_t kNatural[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                              12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                              35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                              58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

constexpr int kLook = 9;  // lookahead bits

// A canonical Huffman table: codes of up to kLook bits resolve in one
// lookup (entry = length << 8 | symbol, 0 = longer code); longer codes by
// comparing with the largest code of each length.
struct Table {
  bool defined = false;
  uint16_t look[1 << kLook];
  int32_t maxcode[17];
  int32_t offset[17];  // symbol index = code + offset[length]
  uint8_t sym[256];
};

bool make_table(const uint8_t counts[16], const uint8_t* syms, int nsyms, Table& t) {
  memset(t.look, 0, sizeof t.look);
  memcpy(t.sym, syms, nsyms);
  int32_t code = 0;
  int k = 0;
  for (int len = 1; len <= 16; ++len) {
    const int cnt = counts[len - 1];
    t.offset[len] = k - code;
    t.maxcode[len] = cnt ? code + cnt - 1 : -1;
    for (int i = 0; i < cnt; ++i, ++code, ++k) {
      if (code >= (1 << len)) return false;  // over-subscribed
      if (len <= kLook) {
        const int shift = kLook - len;
        for (int j = 0; j < (1 << shift); ++j)
          t.look[(code << shift) | j] = (uint16_t)(len << 8 | t.sym[k]);
      }
    }
    code <<= 1;
  }
  t.defined = true;
  return true;
}

struct BitReader {
  const uint8_t* p;
  const uint8_t* end;
  uint64_t buf = 0;  // next bits, MSB first
  int cnt = 0;       // valid bits in buf
  bool at_marker = false;

  // at least 57 valid bits (zeros once a marker or the end is reached)
  void fill() {
    while (cnt <= 56) {
      uint32_t b = 0;
      if (!at_marker && p < end) {
        b = *p;
        if (b != 0xFF) {
          ++p;
        } else if (p + 1 < end && p[1] == 0x00) {
          p += 2;
        } else {
          at_marker = true;
          b = 0;
        }
      }
      buf |= (uint64_t)b << (56 - cnt);
      cnt += 8;
    }
  }
  uint32_t peek(int n) const { return (uint32_t)(buf >> (64 - n)); }
  void skip(int n) {
    buf <<= n;
    cnt -= n;
  }
  int get(int n) {
    if (n == 0) return 0;
    const int v = (int)peek(n);
    skip(n);
    return v;
  }
  int decode(const Table& t) {
    const uint32_t e = t.look[peek(kLook)];
    if (e) {
      skip(e >> 8);
      return e & 0xFF;
    }
    for (int len = kLook + 1; len <= 16; ++len) {
      const int32_t code = (int32_t)peek(len);
      if (code <= t.maxcode[len]) {
        skip(len);
        return t.sym[code + t.offset[len]];
      }
    }
    return -1;
  }
  // restart: drop the buffered bits and pass the next RSTn marker
  bool restart() {
    buf = 0;
    cnt = 0;
    at_marker = false;
    while (p + 1 < end && !(p[0] == 0xFF && p[1] >= 0xD0 && p[1] <= 0xD7)) ++p;
    if (p + 1 >= end) return false;
    p += 2;
    return true;
  }
};

inline int extend(int v, int s) { return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v; }

bool fail(std::string* err, const char* msg) {
  if (err) *err = msg;
  return false;
}

struct Parser {
  const uint8_t* d;
  size_t n;
  kf_jpeg_frame* f;
  int16_t* coef;
  int64_t cap;
  std::string* err;
  Table dc[4], ac[4];
  int ids[KF_JPEG_MAX_COMPONENTS] = {};
  int restart_interval = 0;
  int mcux = 0, mcuy = 0;
  bool jfif = false, adobe = false, have_frame = false, scanned = false;
  int adobe_transform = -1;

  bool frame_header(const uint8_t* s, int len) {
    if (have_frame) return fail(err, "more than one frame header");
    if (len < 6 || s[0] != 8) return fail(err, "only 8-bit samples are supported");
    f->H = s[1] << 8 | s[2];
    f->W = s[3] << 8 | s[4];
    f->nc = s[5];
    if (f->H < 1 || f->W < 1) return fail(err, "image has no rows or columns");
    if ((f->nc != 1 && f->nc != 3) || len < 6 + 3 * f->nc)
      return fail(err, "only 1- and 3-component images are supported");
    f->hmax = f->vmax = 1;
    for (int i = 0; i < f->nc; ++i) {
      kf_jpeg_component& c = f->comp[i];
      ids[i] = s[6 + 3 * i];
      c.h = s[7 + 3 * i] >> 4;
      c.v = s[7 + 3 * i] & 15;
      c.tq = s[8 + 3 * i];
      if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3)
        return fail(err, "bad component sampling or table");
      f->hmax = c.h > f->hmax ? c.h : f->hmax;
      f->vmax = c.v > f->vmax ? c.v : f->vmax;
    }
    mcux = (f->W + 8 * f->hmax - 1) / (8 * f->hmax);
    mcuy = (f->H + 8 * f->vmax - 1) / (8 * f->vmax);
    int64_t b0 = 0;
    for (int i = 0; i < f->nc; ++i) {
      kf_jpeg_component& c = f->comp[i];
      if (f->hmax % c.h || f->vmax % c.v) return fail(err, "non-integral sampling ratios");
      if (f->hmax / c.h == 1 && f->vmax / c.v == 2)
        return fail(err, "1:2 vertical-only chroma sampling (4:4:0) is not supported");
      c.dw = (f->W * c.h + f->hmax - 1) / f->hmax;
      c.dh = (f->H * c.v + f->vmax - 1) / f->vmax;
      c.bw = mcux * c.h;
      c.bh = mcuy * c.v;
      c.block0 = b0;
      b0 += (int64_t)c.bw * c.bh;
    }
    f->n_blocks = b0;
    have_frame = true;
    return true;
  }

  bool quant(const uint8_t* s, int len) {
    int k = 0;
    while (k < len) {
      const int pq = s[k] >> 4, tq = s[k] & 15;
      if (pq > 1 || tq > 3 || k + 1 + 64 * (pq + 1) > len) return fail(err, "bad quantisation table");
      ++k;
      for (int i = 0; i < 64; ++i) {
        f->q[tq][kNatural[i]] = pq ? (uint16_t)(s[k] << 8 | s[k + 1]) : s[k];
        k += pq + 1;
      }
    }
    return true;
  }

  bool huffman(const uint8_t* s, int len) {
    int k = 0;
    while (k < len) {
      if (k + 17 > len) return fail(err, "bad Huffman table");
      const int tc = s[k] >> 4, th = s[k] & 15;
      if (tc > 1 || th > 3) return fail(err, "bad Huffman table class or id");
      int total = 0;
      for (int i = 0; i < 16; ++i) total += s[k + 1 + i];
      if (total > 256 || k + 17 + total > len) return fail(err, "bad Huffman table");
      if (!make_table(s + k + 1, s + k + 17, total, tc ? ac[th] : dc[th]))
        return fail(err, "invalid Huffman code lengths");
      k += 17 + total;
    }
    return true;
  }

  // one scan; returns the position after its entropy-coded data
  bool scan(const uint8_t* s, int len, size_t data_pos, size_t* next) {
    if (!have_frame) return fail(err, "scan before the frame header");
    const int ns = s[0];
    if (ns < 1 || ns > f->nc || len < 4 + 2 * ns) return fail(err, "bad scan header");
    int sc[KF_JPEG_MAX_COMPONENTS], td[KF_JPEG_MAX_COMPONENTS], ta[KF_JPEG_MAX_COMPONENTS];
    for (int i = 0; i < ns; ++i) {
      sc[i] = -1;
      for (int j = 0; j < f->nc; ++j)
        if (ids[j] == s[1 + 2 * i]) sc[i] = j;
      td[i] = s[2 + 2 * i] >> 4;
      ta[i] = s[2 + 2 * i] & 15;
      if (sc[i] < 0 || td[i] > 3 || ta[i] > 3 || !dc[td[i]].defined || !ac[ta[i]].defined)
        return fail(err, "scan names an unknown component or table");
    }
    BitReader br{d + data_pos, d + n};
    int pred[KF_JPEG_MAX_COMPONENTS] = {};
    const bool inter = ns > 1;
    const kf_jpeg_component& c0 = f->comp[sc[0]];
    const int cols = inter ? mcux : (c0.dw + 7) / 8;
    const int rows = inter ? mcuy : (c0.dh + 7) / 8;
    int left = restart_interval;
    for (int my = 0; my < rows; ++my)
      for (int mx = 0; mx < cols; ++mx) {
        if (restart_interval) {
          if (left == 0) {
            if (!br.restart()) return fail(err, "missing restart marker");
            for (int i = 0; i < ns; ++i) pred[i] = 0;
            left = restart_interval;
          }
          --left;
        }
        for (int i = 0; i < ns; ++i) {
          const kf_jpeg_component& c = f->comp[sc[i]];
          const int nby = inter ? c.v : 1, nbx = inter ? c.h : 1;
          for (int by = 0; by < nby; ++by)
            for (int bx = 0; bx < nbx; ++bx) {
              const int gx = inter ? mx * c.h + bx : mx, gy = inter ? my * c.v + by : my;
              int16_t* blk = coef + (c.block0 + (int64_t)gy * c.bw + gx) * 64;
              br.fill();
              const int t = br.decode(dc[td[i]]);
              if (t < 0 || t > 11) return fail(err, "bad DC code");
              if (t) pred[i] += extend(br.get(t), t);
              blk[0] = (int16_t)pred[i];
              for (int k = 1; k < 64; ++k) {
                br.fill();
                const int rs = br.decode(ac[ta[i]]);
                if (rs < 0) return fail(err, "bad AC code");
                const int r = rs >> 4, sz = rs & 15;
                if (sz) {
                  k += r;
                  if (k > 63) return fail(err, "AC run past the end of a block");
                  blk[kNatural[k]] = (int16_t)extend(br.get(sz), sz);
                } else if (r == 15) {
                  k += 15;
                } else {
                  break;
                }
              }
            }
        }
      }
    // the next marker after the entropy-coded data
    const uint8_t* q = br.p;
    while (q + 1 < d + n && !(q[0] == 0xFF && q[1] != 0x00 && !(q[1] >= 0xD0 && q[1] <= 0xD7))) ++q;
    *next = (size_t)(q - d);
    scanned = true;
    return true;
  }

  bool run() {
    if (n < 4 || d[0] != 0xFF || d[1] != 0xD8) return fail(err, "not a JPEG stream");
    size_t pos = 2;
    while (pos + 4 <= n) {
      if (d[pos] != 0xFF) return fail(err, "marker expected");
      const int m = d[pos + 1];
      if (m == 0xFF) {
        ++pos;
        continue;
      }
      if (m == 0xD9) break;
      const int len = d[pos + 2] << 8 | d[pos + 3];
      if (len < 2 || pos + 2 + (size_t)len > n) return fail(err, "truncated segment");
      const uint8_t* s = d + pos + 4;
      const int sl = len - 2;
      if (m == 0xC0 || m == 0xC1) {
        if (!frame_header(s, sl)) return false;
        if (!coef) return true;
        if (cap < f->n_blocks) return fail(err, "coefficient buffer too small");
        memset(coef, 0, (size_t)f->n_blocks * 64 * sizeof(int16_t));
      } else if (m == 0xC2 || m == 0xC3 || (m >= 0xC5 && m <= 0xCF && m != 0xC8 && m != 0xCC)) {
        return fail(err, "progressive, lossless, hierarchical and arithmetic-coded JPEG are not "
                         "supported");
      } else if (m == 0xDB) {
        if (!quant(s, sl)) return false;
      } else if (m == 0xC4) {
        if (!huffman(s, sl)) return false;
      } else if (m == 0xDD) {
        if (sl < 2) return fail(err, "bad restart interval");
        restart_interval = s[0] << 8 | s[1];
      } else if (m == 0xE0) {
        if (sl >= 5 && !memcmp(s, "JFIF\0", 5)) jfif = true;
      } else if (m == 0xEE) {
        if (sl >= 12 && !memcmp(s, "Adobe", 5)) {
          adobe = true;
          adobe_transform = s[11];
        }
      } else if (m == 0xDA) {
        size_t next = 0;
        if (!scan(s, sl, pos + 2 + len, &next)) return false;
        pos = next;
        continue;
      }
      pos += 2 + len;
    }
    if (!have_frame) return fail(err, "no frame header");
    if (coef && !scanned) return fail(err, "no scan");
    return true;
  }
};

}  // namespace

int kf_jpeg_decode_entropy(const uint8_t* data, size_t n, kf_jpeg_frame* f, int16_t* coef,
                           int64_t capacity, std::string* err) {
  if (!data || !f) {
    if (err) *err = "null argument";
    return -1;
  }
  *f = kf_jpeg_frame{};
  Parser p{data, n, f, coef, capacity, err};
  if (!p.run()) return -1;
  // YCbCr unless RGB: a JFIF marker means YCbCr, else an Adobe marker's
  // transform 0 means RGB, else component ids 'R', 'G', 'B'
  f->rgb = f->nc == 3 && !p.jfif &&
           (p.adobe ? p.adobe_transform == 0 : p.ids[0] == 'R' && p.ids[1] == 'G' && p.ids[2] == 'B');
  return 0;
}
