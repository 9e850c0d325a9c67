// Dataset-format IO of the reference's frame loop, without OpenImageIO:
//   - OpenEXR scanline images in, as interleaved RGB float
//     (opencl/bmfr.cpp:145-172: color, shading_normal, world_position and
//     albedo NN.exr, 3 channels, converted to float);
//   - PNG (8-bit RGB) and EXR (FLOAT) out with the image-width crop
//     (opencl/bmfr.cpp:519-547);
//   - the dataset's camera_matrices.h: per-frame 4x4 view-projection
//     matrices, pixel offsets and the two reprojection limits
//     (opencl/bmfr.cpp:43-47, 226-227, 440-444), parsed at run time instead
//     of being compiled in.
// Host code only; errors are status codes with a message (bmfr_last_error).
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "bmfr_hip.h"
#include "host_error.h"
#include "surface.h"

namespace {

using bmfr_host::fail;

bool read_file(const char *path, std::vector<uint8_t> &out) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        return false;
    }
    out.resize((size_t)n);
    const size_t got = n ? std::fread(out.data(), 1, (size_t)n, f) : 0;
    std::fclose(f);
    return got == (size_t)n;
}

bool write_file(const char *path, const std::vector<uint8_t> &data) {
    FILE *f = std::fopen(path, "wb");
    if (!f) return false;
    const size_t put = data.empty() ? 0 : std::fwrite(data.data(), 1, data.size(), f);
    const bool ok = put == data.size();
    return std::fclose(f) == 0 && ok;
}

// ---- OpenEXR (scanline, single part) -----------------------------------------
// Layout: magic 76 2f 31 01, version 2 (flags: tiled 0x200, multipart
// 0x1000, deep 0x800), attributes (name, type, int32 size, value) up to an
// empty name, one uint64 chunk offset per chunk, then chunks of
// (int32 y, int32 size, data).  A chunk's raw data is, per scanline, each
// channel's row in channel-list (alphabetical) order, little endian.
enum ExrCompression { kNone = 0, kRle = 1, kZips = 2, kZip = 3 };
enum ExrPixel { kUint = 0, kHalf = 1, kFloat = 2 };

struct ExrChannel {
    std::string name;
    int32_t type = kFloat;
    int32_t xs = 1, ys = 1;
};

struct ExrImage {
    int32_t xmin = 0, ymin = 0, xmax = -1, ymax = -1;
    int compression = kNone;
    std::vector<ExrChannel> channels;
    size_t header_end = 0;
    // In 64 bits: a corrupted window must not overflow before it is checked.
    int64_t width() const { return (int64_t)xmax - xmin + 1; }
    int64_t height() const { return (int64_t)ymax - ymin + 1; }
};

constexpr int64_t kMaxExrSide = 65536;   // larger windows are refused as malformed

struct Reader {
    const std::vector<uint8_t> &b;
    size_t p = 0;
    bool ok = true;
    explicit Reader(const std::vector<uint8_t> &buf) : b(buf) {}
    // n more bytes from p are inside the buffer (no wrap for any p)
    bool fits(size_t n) const { return p <= b.size() && b.size() - p >= n; }
    template <typename T>
    T get() {
        T v{};
        if (!fits(sizeof(T))) {
            ok = false;
            return v;
        }
        std::memcpy(&v, b.data() + p, sizeof(T));
        p += sizeof(T);
        return v;
    }
    std::string str() {
        std::string s;
        while (p < b.size() && b[p]) s.push_back((char)b[p++]);
        if (p >= b.size()) ok = false;
        else ++p;
        return s;
    }
};

int32_t parse_exr_header(const std::vector<uint8_t> &buf, ExrImage &im, const char *path) {
    Reader r(buf);
    const uint32_t magic = r.get<uint32_t>();
    const uint32_t version = r.get<uint32_t>();
    if (!r.ok || magic != 20000630u) return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: not an OpenEXR file", path);
    if ((version & 0xff) != 2 || (version & (0x200 | 0x800 | 0x1000)))
        return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: only single-part scanline OpenEXR is supported", path);
    bool have_channels = false, have_window = false;
    while (r.ok) {
        const std::string name = r.str();
        if (!r.ok) break;
        if (name.empty()) break;
        const std::string type = r.str();
        const int32_t size = r.get<int32_t>();
        if (!r.ok || size < 0 || !r.fits((size_t)size)) {
            r.ok = false;
            break;
        }
        const size_t end = r.p + (size_t)size;
        if (name == "channels" && type == "chlist") {
            while (r.ok && r.p < end) {
                ExrChannel c;
                c.name = r.str();
                if (!r.ok || c.name.empty()) break;
                c.type = r.get<int32_t>();
                (void)r.get<uint8_t>();   // pLinear
                r.p += 3;                 // reserved
                c.xs = r.get<int32_t>();
                c.ys = r.get<int32_t>();
                im.channels.push_back(c);
            }
            have_channels = true;
        } else if (name == "compression" && type == "compression") {
            im.compression = r.get<uint8_t>();
        } else if (name == "dataWindow" && type == "box2i") {
            im.xmin = r.get<int32_t>();
            im.ymin = r.get<int32_t>();
            im.xmax = r.get<int32_t>();
            im.ymax = r.get<int32_t>();
            have_window = true;
        }
        r.p = end;
    }
    if (!r.ok || !have_channels || !have_window || im.width() <= 0 || im.height() <= 0 ||
        im.width() > kMaxExrSide || im.height() > kMaxExrSide)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: malformed OpenEXR header", path);
    for (const ExrChannel &c : im.channels) {
        if (c.xs != 1 || c.ys != 1 || c.type < kUint || c.type > kFloat)
            return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: channel %s is subsampled or of an unknown type", path,
                        c.name.c_str());
    }
    if (im.compression > kZip)
        return fail(BMFR_ERR_INVALID_ARGUMENT,
                    "%s: OpenEXR compression %d is not supported (NONE, RLE, ZIPS and ZIP are)", path,
                    im.compression);
    im.header_end = r.p;
    return BMFR_OK;
}

float half_to_float(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31;
    const int e = (h >> 10) & 0x1f;
    const uint32_t m = h & 0x3ff;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = s;
        } else {  // subnormal: m * 2^-24
            const float f = std::ldexp((float)m, -24);
            std::memcpy(&bits, &f, 4);
            bits |= s;
        }
    } else if (e == 31) {
        bits = s | 0x7f800000u | (m << 13);
    } else {
        bits = s | ((uint32_t)(e - 15 + 127) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// Undo the ZIP/RLE pre-pass: byte predictor (d[i] = t[i] - t[i-1] + 128),
// then the split of even and odd bytes into two halves.
void exr_unpredict(std::vector<uint8_t> &t) {
    for (size_t i = 1; i < t.size(); ++i) t[i] = (uint8_t)(t[i - 1] + t[i] - 128);
    std::vector<uint8_t> out(t.size());
    const size_t half = (t.size() + 1) / 2;
    for (size_t i = 0; i < t.size(); ++i) out[i] = (i & 1) ? t[half + i / 2] : t[i / 2];
    t.swap(out);
}

void exr_predict(std::vector<uint8_t> &raw) {
    std::vector<uint8_t> t(raw.size());
    const size_t half = (raw.size() + 1) / 2;
    for (size_t i = 0; i < raw.size(); ++i) {
        if (i & 1) t[half + i / 2] = raw[i];
        else t[i / 2] = raw[i];
    }
    for (size_t i = t.size(); i-- > 1;) t[i] = (uint8_t)(t[i] - t[i - 1] + 128);
    raw.swap(t);
}

bool rle_decode(const uint8_t *in, size_t n, std::vector<uint8_t> &out, size_t expect) {
    out.clear();
    size_t p = 0;
    while (p < n) {
        const int8_t c = (int8_t)in[p++];
        if (c < 0) {
            const size_t k = (size_t)(-(int)c);
            if (p + k > n) return false;
            out.insert(out.end(), in + p, in + p + k);
            p += k;
        } else {
            if (p >= n) return false;
            out.insert(out.end(), (size_t)c + 1, in[p++]);
        }
        if (out.size() > expect) return false;
    }
    return out.size() == expect;
}

int lines_per_chunk(int compression) { return compression == kZip ? 16 : 1; }

size_t pixel_bytes(int32_t type) { return type == kHalf ? 2 : 4; }

int32_t load_exr(const char *path, std::vector<uint8_t> &buf, ExrImage &im) {
    if (!path) return fail(BMFR_ERR_INVALID_ARGUMENT, "path is null");
    if (!read_file(path, buf)) return fail(BMFR_ERR_INVALID_ARGUMENT, "cannot read %s", path);
    return parse_exr_header(buf, im, path);
}

// Index of channel `want` (or -1).
int find_channel(const ExrImage &im, const char *want) {
    for (size_t i = 0; i < im.channels.size(); ++i)
        if (im.channels[i].name == want) return (int)i;
    return -1;
}

// ---- PNG ---------------------------------------------------------------------
void put_be32(std::vector<uint8_t> &v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void png_chunk(std::vector<uint8_t> &out, const char type[4], const std::vector<uint8_t> &data) {
    put_be32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    const uLong crc = crc32(0L, out.data() + start, (uInt)(out.size() - start));
    put_be32(out, (uint32_t)crc);
}

// ---- camera_matrices.h ---------------------------------------------------------
// The dataset's header declares (as C arrays or macros):
//   camera_matrices[FRAMES][4][4], pixel_offsets[FRAMES][2],
//   position_limit_squared, normal_limit_squared.
// Every declaration `... name[..]... = { numbers };` or `... name = number;`
// or `#define name number` is collected by name, numbers in source order.
struct Token {
    enum Kind { kIdent, kNumber, kPunct } kind;
    std::string text;
};

std::vector<Token> tokenize(const std::string &s) {
    std::vector<Token> t;
    size_t i = 0;
    while (i < s.size()) {
        const char c = s[i];
        if (std::isspace((unsigned char)c)) {
            ++i;
        } else if (c == '/' && i + 1 < s.size() && s[i + 1] == '/') {
            while (i < s.size() && s[i] != '\n') ++i;
        } else if (c == '/' && i + 1 < s.size() && s[i + 1] == '*') {
            const size_t e = s.find("*/", i + 2);
            i = e == std::string::npos ? s.size() : e + 2;
        } else if (std::isalpha((unsigned char)c) || c == '_' || c == '#') {
            size_t j = i + 1;
            while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_')) ++j;
            t.push_back({Token::kIdent, s.substr(i, j - i)});
            i = j;
        } else if (std::isdigit((unsigned char)c) || (c == '.' && i + 1 < s.size() && std::isdigit((unsigned char)s[i + 1]))) {
            size_t j = i;
            while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '.' ||
                                    ((s[j] == '-' || s[j] == '+') && (s[j - 1] == 'e' || s[j - 1] == 'E'))))
                ++j;
            t.push_back({Token::kNumber, s.substr(i, j - i)});
            i = j;
        } else {
            t.push_back({Token::kPunct, std::string(1, c)});
            ++i;
        }
    }
    return t;
}

bool number_value(const std::string &text, double &v) {
    std::string x = text;
    while (!x.empty() && (x.back() == 'f' || x.back() == 'F' || x.back() == 'l' || x.back() == 'L' ||
                          x.back() == 'u' || x.back() == 'U'))
        x.pop_back();
    char *end = nullptr;
    v = std::strtod(x.c_str(), &end);
    return end && *end == '\0' && !x.empty();
}

int32_t parse_camera_header(const std::string &src, std::unordered_map<std::string, std::vector<double>> &vars) {
    const std::vector<Token> t = tokenize(src);
    for (size_t i = 0; i < t.size(); ++i) {
        if (t[i].kind == Token::kIdent && t[i].text == "#define" && i + 2 < t.size() &&
            t[i + 1].kind == Token::kIdent) {
            // #define NAME [-]number
            size_t j = i + 2;
            double sign = 1.0;
            if (t[j].kind == Token::kPunct && t[j].text == "-" && j + 1 < t.size()) {
                sign = -1.0;
                ++j;
            }
            double v;
            if (t[j].kind == Token::kNumber && number_value(t[j].text, v)) vars[t[i + 1].text] = {sign * v};
            continue;
        }
        if (!(t[i].kind == Token::kPunct && t[i].text == "=")) continue;
        // the declared name: the last identifier before '=' outside brackets
        std::string name;
        int depth = 0;
        for (size_t k = i; k-- > 0;) {
            if (t[k].kind == Token::kPunct && t[k].text == "]") ++depth;
            else if (t[k].kind == Token::kPunct && t[k].text == "[") --depth;
            else if (depth == 0 && t[k].kind == Token::kIdent) {
                name = t[k].text;
                break;
            } else if (depth == 0 && t[k].kind == Token::kPunct && (t[k].text == ";" || t[k].text == "}")) {
                break;
            }
        }
        if (name.empty()) continue;
        std::vector<double> vals;
        size_t j = i + 1;
        double sign = 1.0;
        for (; j < t.size() && !(t[j].kind == Token::kPunct && t[j].text == ";"); ++j) {
            if (t[j].kind == Token::kPunct && t[j].text == "-") {
                sign = -sign;
            } else if (t[j].kind == Token::kNumber) {
                double v;
                if (!number_value(t[j].text, v))
                    return fail(BMFR_ERR_INVALID_ARGUMENT, "camera header: bad number '%s' in %s",
                                t[j].text.c_str(), name.c_str());
                vals.push_back(sign * v);
                sign = 1.0;
            } else if (t[j].kind == Token::kPunct && (t[j].text == "," || t[j].text == "{" || t[j].text == "}")) {
                sign = 1.0;
            }
        }
        vars[name] = vals;
        i = j;
    }
    return BMFR_OK;
}

}  // namespace

extern "C" {

int32_t bmfr_exr_info(const char *path, int32_t *width, int32_t *height, int32_t *channels) {
    std::vector<uint8_t> buf;
    ExrImage im;
    if (int32_t st = load_exr(path, buf, im)) return st;
    if (width) *width = (int32_t)im.width();
    if (height) *height = (int32_t)im.height();
    if (channels) *channels = (int32_t)im.channels.size();
    return BMFR_OK;
}

int32_t bmfr_exr_read_rgb(const char *path, int32_t width, int32_t height, float *rgb) {
    if (!rgb) return fail(BMFR_ERR_INVALID_ARGUMENT, "output is null");
    std::vector<uint8_t> buf;
    ExrImage im;
    if (int32_t st = load_exr(path, buf, im)) return st;
    const int W = (int)im.width(), H = (int)im.height();
    if (W != width || H != height)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "%s is %dx%d, expected %dx%d", path, W, H, width, height);
    // R, G, B by name; a single luminance channel Y fills all three
    int src[3] = {find_channel(im, "R"), find_channel(im, "G"), find_channel(im, "B")};
    if (src[0] < 0 || src[1] < 0 || src[2] < 0) {
        const int y = find_channel(im, "Y");
        if (y < 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "%s has no R, G, B (or Y) channels", path);
        src[0] = src[1] = src[2] = y;
    }
    size_t row_bytes = 0;
    std::vector<size_t> ch_off;
    for (const ExrChannel &c : im.channels) {
        ch_off.push_back(row_bytes);
        row_bytes += pixel_bytes(c.type) * (size_t)W;
    }
    const int lpc = lines_per_chunk(im.compression);
    const int chunks = (H + lpc - 1) / lpc;
    Reader r(buf);
    r.p = im.header_end;
    std::vector<uint64_t> offsets((size_t)chunks);
    for (int c = 0; c < chunks; ++c) offsets[c] = r.get<uint64_t>();
    if (!r.ok) return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: truncated chunk table", path);
    std::vector<uint8_t> raw, tmp;
    std::vector<uint8_t> filled((size_t)H, 0);   // every line exactly once
    for (int c = 0; c < chunks; ++c) {
        if (offsets[c] > (uint64_t)buf.size())
            return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: chunk %d offset beyond the file", path, c);
        Reader cr(buf);
        cr.p = (size_t)offsets[c];
        const int32_t y0 = cr.get<int32_t>();
        const int32_t size = cr.get<int32_t>();
        if (!cr.ok || size < 0 || !cr.fits((size_t)size))
            return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: bad chunk %d", path, c);
        const int64_t line0_64 = (int64_t)y0 - im.ymin;
        if (line0_64 < 0 || line0_64 >= H)
            return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: chunk %d starts outside", path, c);
        const int line0 = (int)line0_64;
        const int lines = std::min(lpc, H - line0);
        for (int l = 0; l < lines; ++l) {
            if (filled[(size_t)(line0 + l)])
                return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: line %d appears in two chunks", path, line0 + l);
            filled[(size_t)(line0 + l)] = 1;
        }
        const size_t expect = row_bytes * (size_t)lines;
        const uint8_t *data = buf.data() + cr.p;
        if ((size_t)size == expect || im.compression == kNone) {
            if ((size_t)size != expect) return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: chunk %d has %d bytes", path, c, size);
            raw.assign(data, data + expect);
        } else if (im.compression == kRle) {
            if (!rle_decode(data, (size_t)size, raw, expect))
                return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: bad RLE chunk %d", path, c);
            exr_unpredict(raw);
        } else {
            // deflate expands at most about 1032:1; a larger claim is malformed
            // (and would otherwise size the buffer from the header alone)
            if (expect / 1032 > (size_t)size + 64)
                return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: ZIP chunk %d claims %zu bytes from %d", path, c, expect, size);
            raw.resize(expect);
            uLongf got = (uLongf)expect;
            if (uncompress(raw.data(), &got, data, (uLong)size) != Z_OK || got != expect)
                return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: bad ZIP chunk %d", path, c);
            exr_unpredict(raw);
        }
        for (int l = 0; l < lines; ++l) {
            const uint8_t *row = raw.data() + row_bytes * (size_t)l;
            float *dst = rgb + ((size_t)(line0 + l) * W) * 3;
            for (int k = 0; k < 3; ++k) {
                const ExrChannel &ch = im.channels[src[k]];
                const uint8_t *p = row + ch_off[src[k]];
                for (int x = 0; x < W; ++x) {
                    float v;
                    if (ch.type == kHalf) {
                        uint16_t h;
                        std::memcpy(&h, p + 2 * x, 2);
                        v = half_to_float(h);
                    } else if (ch.type == kFloat) {
                        std::memcpy(&v, p + 4 * x, 4);
                    } else {
                        uint32_t u;
                        std::memcpy(&u, p + 4 * x, 4);
                        v = (float)u;
                    }
                    dst[3 * x + k] = v;
                }
            }
        }
    }
    for (int y = 0; y < H; ++y)
        if (!filled[(size_t)y]) return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: line %d is in no chunk", path, y);
    return BMFR_OK;
}

int32_t bmfr_exr_write_rgb(const char *path, int32_t width, int32_t height, const float *rgb,
                           int64_t row_stride, int32_t compression) {
    if (!path || !rgb || width <= 0 || height <= 0 || row_stride < width)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "bad EXR write arguments");
    if (compression != kNone && compression != kZip)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "EXR output compression %d: use 0 (NONE) or 3 (ZIP)", compression);
    std::vector<uint8_t> out;
    auto u32 = [&](uint32_t x) { out.insert(out.end(), (uint8_t *)&x, (uint8_t *)&x + 4); };
    auto s = [&](const char *x) { out.insert(out.end(), x, x + std::strlen(x) + 1); };
    auto attr = [&](const char *name, const char *type, const std::vector<uint8_t> &v) {
        s(name);
        s(type);
        u32((uint32_t)v.size());
        out.insert(out.end(), v.begin(), v.end());
    };
    auto bytes = [](const void *p, size_t n) {
        return std::vector<uint8_t>((const uint8_t *)p, (const uint8_t *)p + n);
    };
    u32(20000630u);
    u32(2u);
    std::vector<uint8_t> chl;
    for (const char *n : {"B", "G", "R"}) {
        chl.insert(chl.end(), n, n + 2);
        const int32_t fields[4] = {kFloat, 0, 1, 1};   // type, pLinear + reserved, xs, ys
        const uint8_t *fp = (const uint8_t *)fields;
        chl.insert(chl.end(), fp, fp + 16);
    }
    chl.push_back(0);
    attr("channels", "chlist", chl);
    const uint8_t comp = (uint8_t)compression;
    attr("compression", "compression", bytes(&comp, 1));
    const int32_t win[4] = {0, 0, width - 1, height - 1};
    attr("dataWindow", "box2i", bytes(win, 16));
    attr("displayWindow", "box2i", bytes(win, 16));
    const uint8_t order = 0;
    attr("lineOrder", "lineOrder", bytes(&order, 1));
    const float par = 1.f;
    attr("pixelAspectRatio", "float", bytes(&par, 4));
    const float swc[2] = {0.f, 0.f};
    attr("screenWindowCenter", "v2f", bytes(swc, 8));
    attr("screenWindowWidth", "float", bytes(&par, 4));
    out.push_back(0);
    const int lpc = lines_per_chunk(compression);
    const int chunks = (height + lpc - 1) / lpc;
    const size_t table = out.size();
    out.resize(out.size() + 8 * (size_t)chunks);
    const size_t row_bytes = 12 * (size_t)width;
    for (int c = 0; c < chunks; ++c) {
        const int line0 = c * lpc, lines = std::min(lpc, height - line0);
        std::vector<uint8_t> raw(row_bytes * (size_t)lines);
        for (int l = 0; l < lines; ++l) {
            const float *src = rgb + (size_t)(line0 + l) * (size_t)row_stride * 3;
            for (int k = 0; k < 3; ++k) {   // B, G, R
                float *dst = (float *)(raw.data() + row_bytes * (size_t)l + 4 * (size_t)width * k);
                for (int x = 0; x < width; ++x) dst[x] = src[3 * x + (2 - k)];
            }
        }
        std::vector<uint8_t> data = raw;
        if (compression == kZip) {
            std::vector<uint8_t> t = raw;
            exr_predict(t);
            uLongf n = compressBound((uLong)t.size());
            std::vector<uint8_t> z(n);
            if (compress2(z.data(), &n, t.data(), (uLong)t.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
                return fail(BMFR_ERR_INVALID_ARGUMENT, "zlib failed");
            if (n < raw.size()) data.assign(z.begin(), z.begin() + n);
        }
        const uint64_t off = out.size();
        std::memcpy(out.data() + table + 8 * (size_t)c, &off, 8);
        u32((uint32_t)line0);
        u32((uint32_t)data.size());
        out.insert(out.end(), data.begin(), data.end());
    }
    if (!write_file(path, out)) return fail(BMFR_ERR_INVALID_ARGUMENT, "cannot write %s", path);
    return BMFR_OK;
}

int32_t bmfr_png_write_rgb(const char *path, int32_t width, int32_t height, const float *rgb, int64_t row_stride) {
    if (!path || !rgb || width <= 0 || height <= 0 || row_stride < width)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "bad PNG write arguments");
    // rows: filter byte 0, then 8-bit RGB = round(clamp(v, 0, 1) * 255)
    std::vector<uint8_t> rows((size_t)height * (1 + 3 * (size_t)width));
    for (int y = 0; y < height; ++y) {
        uint8_t *d = rows.data() + (size_t)y * (1 + 3 * (size_t)width);
        d[0] = 0;
        const float *src = rgb + (size_t)y * (size_t)row_stride * 3;
        for (int i = 0; i < 3 * width; ++i) d[1 + i] = bmfr::quantize_u8(src[i]);   // NaN -> 0
    }
    uLongf n = compressBound((uLong)rows.size());
    std::vector<uint8_t> z(n);
    if (compress2(z.data(), &n, rows.data(), (uLong)rows.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
        return fail(BMFR_ERR_INVALID_ARGUMENT, "zlib failed");
    z.resize(n);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)width);
    put_be32(ihdr, (uint32_t)height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});   // 8-bit, RGB, deflate, filter 0, no interlace
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", {});
    if (!write_file(path, out)) return fail(BMFR_ERR_INVALID_ARGUMENT, "cannot write %s", path);
    return BMFR_OK;
}

int32_t bmfr_camera_matrices_load(const char *path, int32_t frames, float *matrices, float *offsets,
                                  float *position_limit_squared, float *normal_limit_squared) {
    if (!path || frames <= 0) return fail(BMFR_ERR_INVALID_ARGUMENT, "bad camera header arguments");
    std::vector<uint8_t> buf;
    if (!read_file(path, buf)) return fail(BMFR_ERR_INVALID_ARGUMENT, "cannot read %s", path);
    std::unordered_map<std::string, std::vector<double>> vars;
    if (int32_t st = parse_camera_header(std::string(buf.begin(), buf.end()), vars)) return st;
    auto take = [&](const char *name, size_t need, float *dst) -> int32_t {
        if (!dst) return BMFR_OK;
        auto it = vars.find(name);
        if (it == vars.end()) return fail(BMFR_ERR_INVALID_ARGUMENT, "%s declares no %s", path, name);
        if (it->second.size() < need)
            return fail(BMFR_ERR_INVALID_ARGUMENT, "%s: %s has %zu values, %zu needed", path, name,
                        it->second.size(), need);
        for (size_t i = 0; i < need; ++i) dst[i] = (float)it->second[i];
        return BMFR_OK;
    };
    if (int32_t st = take("camera_matrices", 16 * (size_t)frames, matrices)) return st;
    if (int32_t st = take("pixel_offsets", 2 * (size_t)frames, offsets)) return st;
    if (int32_t st = take("position_limit_squared", 1, position_limit_squared)) return st;
    return take("normal_limit_squared", 1, normal_limit_squared);
}

}  // extern "C"
