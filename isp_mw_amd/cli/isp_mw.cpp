// isp_mw - command-line front end of the BM3D and NL-means filters, with the
// flags and file behaviour of ISP_MW's `--bm3d` and `--nlm` (Filtering,
// source/ISP_MW.cpp:201-289; FilterIO, include/Filter.h:220-340; BM3D_IO,
// include/BM3D.h:363-526; NLMeans_IO, include/NLMeans.h:62-171), on
// libisp_bm3d.so and without OpenCV.
//
//   isp_mw --bm3d <input> [-T tag] [-F format] [-P profile] [--ref path]
//          [-S sigma]... [-BS1 n] [-BSP1 n] [-GS1 n] [-MR1 n] [-MS1 n] [-TH1 x] [-LD x]
//          [-BS2 n] [-BSP2 n] [-GS2 n] [-MR2 n] [-MS2 n] [-TH2 x]
//          [-D device] [--precision fp32|fp32fast|fp32mfma|fp16|bf16] [--print-params]
//   isp_mw --nlm|--nlmeans|--nonlocalmeans <input> [-T tag] [-F format] [--ref path]
//          [-C true|false] [-S sigma] [-H strength] [-BS n] [-BSP n] [-GS n] [-MR n] [-MS n]
//          [-TH x] [-D device] [--precision fp32|fp32fast|fp32mfma|fp16|bf16] [--print-params]
//
// BM3D as the reference: the profile defaults to "fast"; -S may repeat (one
// value per plane, the last one repeated to three) and sets both steps; thMSE
// follows thMSE_Default unless -TH1 / -TH2 give it; the output is
// <input dir>/<input stem><tag><format>, tag ".BM3D", format ".png".
// NL-means as the reference: strength and thMSE follow the final sigma and
// correction (5 / 50 sigma, 1.5 / 25 sigma without correction) unless -H /
// -TH give them; tag ".NLMeans".  The filter name is matched case-blind.
// 8-bit images become 16-bit RGB frames (value * 257, grey replicated to
// R, G, B, include/ImageIO.h ImageReader) filtered by the integer Frame path
// and written as 8-bit RGB with RoundDiv(value, 257) (ImageWriter).
// Files: PNG (8/16-bit grey, grey+alpha, RGB, RGBA, palette; not interlaced),
// PPM/PGM (P5/P6, 8/16-bit) and PFM (float RGB or grey, filtered by BM3D as a
// float frame in [0, 1] and written as PFM; NL-means filters integer frames
// only, as the reference's Frame).  16-bit PNG / PNM samples are kept as
// 16-bit values (the reference's OpenCV reader reduces them to 8 bits).
// The extra flags: -D the HIP device, --precision the transform precision
// (NL-means: FP64 distances for fp32, FP32 distances otherwise),
// --print-params the parsed parameters as JSON without filtering, --copy the
// file conversion alone (read, 16-bit frame, write) without filtering.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "isp_bm3d.h"
#include "isp_nlm.h"

namespace {

// largest image side a file may declare (the library's own limits are
// checked by the library: width < 65536, height < 32768)
constexpr uint32_t kMaxSide = 65535;

struct Image {
    int w = 0, h = 0;
    int channels = 3;         // after loading: always 3 (R, G, B)
    int depth = 8;            // 8 or 16 for integer files; 32 for float
    bool grey_source = false;
    std::vector<int32_t> px;  // planar R, G, B (integer files)
    std::vector<float> fx;    // planar R, G, B (PFM)
};

[[noreturn]] void die(const std::string &m)
{
    throw std::runtime_error(m);
}

int precision_code(const std::string &p);

std::vector<unsigned char> read_file(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) die("could not open " + path);
    return std::vector<unsigned char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

void write_file(const std::string &path, const std::vector<unsigned char> &data)
{
    std::ofstream f(path, std::ios::binary);
    if (!f) die("could not write " + path);
    f.write(reinterpret_cast<const char *>(data.data()), (std::streamsize)data.size());
    if (!f) die("could not write " + path);
}

uint32_t be32(const unsigned char *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

void put32(std::vector<unsigned char> &o, uint32_t v)
{
    for (int s = 24; s >= 0; s -= 8) o.push_back((unsigned char)(v >> s));
}

// ---- PNG (ISO/IEC 15948): chunks, zlib stream, per-row filters 0..4 ----
Image read_png(const std::vector<unsigned char> &d)
{
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (d.size() < 8 || std::memcmp(d.data(), sig, 8)) die("not a PNG file");
    size_t p = 8;
    int w = 0, h = 0, bd = 0, ct = -1, il = 0;
    std::vector<unsigned char> idat, plte;
    while (p + 8 <= d.size()) {
        const uint32_t len = be32(&d[p]);
        const std::string type(reinterpret_cast<const char *>(&d[p + 4]), 4);
        if (p + 12 + (size_t)len > d.size()) die("truncated PNG chunk " + type);
        const unsigned char *c = &d[p + 8];
        if (type == "IHDR") {
            if (len != 13) die("PNG IHDR chunk of " + std::to_string(len) + " bytes (13 expected)");
            const uint32_t uw = be32(c), uh = be32(c + 4);
            if (uw == 0 || uh == 0 || uw > kMaxSide || uh > kMaxSide) die("PNG dimensions out of range");
            w = (int)uw;
            h = (int)uh;
            bd = c[8];
            ct = c[9];
            il = c[12];
        } else if (type == "PLTE") {
            plte.assign(c, c + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), c, c + len);
        } else if (type == "IEND") {
            break;
        }
        p += 12 + len;
    }
    if (w <= 0 || h <= 0) die("PNG without a valid IHDR");
    if (il != 0) die("interlaced PNG is not supported");
    int spp;  // samples per pixel
    switch (ct) {
    case 0: spp = 1; break;
    case 2: spp = 3; break;
    case 3: spp = 1; break;
    case 4: spp = 2; break;
    case 6: spp = 4; break;
    default: die("unsupported PNG colour type " + std::to_string(ct));
    }
    if (!((bd == 8 || bd == 16) && (ct != 3 || bd == 8))) die("unsupported PNG bit depth " + std::to_string(bd));
    const size_t bpp = (size_t)spp * (bd / 8), stride = (size_t)w * bpp;
    std::vector<unsigned char> raw((stride + 1) * (size_t)h);
    uLongf rl = (uLongf)raw.size();
    if (uncompress(raw.data(), &rl, idat.data(), (uLong)idat.size()) != Z_OK || rl != raw.size())
        die("corrupt PNG image data");
    std::vector<unsigned char> cur(stride), prev(stride, 0);
    Image im;
    im.w = w;
    im.h = h;
    im.depth = bd;
    im.grey_source = ct == 0 || ct == 4;
    const size_t np = (size_t)w * h;
    im.px.assign(3 * np, 0);
    for (int y = 0; y < h; ++y) {
        const unsigned char *r = &raw[(stride + 1) * (size_t)y];
        const int ft = r[0];
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int v = r[1 + i];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) / 2; break;
            case 4: {
                const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                break;
            }
            default: die("bad PNG row filter");
            }
            cur[i] = (unsigned char)v;
        }
        for (int x = 0; x < w; ++x) {
            auto sample = [&](int s) -> int {
                const size_t o = (size_t)x * bpp + (size_t)s * (bd / 8);
                return bd == 8 ? cur[o] : (cur[o] << 8 | cur[o + 1]);
            };
            int rgb[3];
            if (ct == 3) {
                const int k = sample(0);
                if ((size_t)(3 * k + 2) >= plte.size()) die("PNG palette index out of range");
                for (int q = 0; q < 3; ++q) rgb[q] = plte[3 * k + q];
            } else if (spp >= 3) {
                for (int q = 0; q < 3; ++q) rgb[q] = sample(q);
            } else {
                rgb[0] = rgb[1] = rgb[2] = sample(0);
            }
            for (int q = 0; q < 3; ++q) im.px[q * np + (size_t)y * w + x] = rgb[q];
        }
        std::swap(cur, prev);
    }
    return im;
}

void chunk(std::vector<unsigned char> &o, const char *type, const std::vector<unsigned char> &data)
{
    put32(o, (uint32_t)data.size());
    const size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    put32(o, (uint32_t)crc32(0, &o[start], (uInt)(o.size() - start)));
}

// RGB, 8 or 16 bits, rows with filter 1 (Sub)
std::vector<unsigned char> write_png(const Image &im, int depth)
{
    const size_t np = (size_t)im.w * im.h, bpp = 3 * (size_t)(depth / 8), stride = (size_t)im.w * bpp;
    std::vector<unsigned char> raw((stride + 1) * (size_t)im.h), row(stride);
    for (int y = 0; y < im.h; ++y) {
        for (int x = 0; x < im.w; ++x)
            for (int q = 0; q < 3; ++q) {
                const int v = im.px[q * np + (size_t)y * im.w + x];
                const size_t o = (size_t)x * bpp + (size_t)q * (depth / 8);
                if (depth == 8) row[o] = (unsigned char)v;
                else {
                    row[o] = (unsigned char)(v >> 8);
                    row[o + 1] = (unsigned char)v;
                }
            }
        unsigned char *r = &raw[(stride + 1) * (size_t)y];
        r[0] = 1;
        for (size_t i = 0; i < stride; ++i) r[1 + i] = (unsigned char)(row[i] - (i >= bpp ? row[i - bpp] : 0));
    }
    uLongf cl = compressBound((uLong)raw.size());
    std::vector<unsigned char> z(cl);
    if (compress2(z.data(), &cl, raw.data(), (uLong)raw.size(), 6) != Z_OK) die("deflate failed");
    z.resize(cl);
    std::vector<unsigned char> o = {137, 80, 78, 71, 13, 10, 26, 10}, hdr;
    put32(hdr, (uint32_t)im.w);
    put32(hdr, (uint32_t)im.h);
    hdr.push_back((unsigned char)depth);
    hdr.push_back(2);  // RGB
    hdr.push_back(0);
    hdr.push_back(0);
    hdr.push_back(0);
    chunk(o, "IHDR", hdr);
    chunk(o, "IDAT", z);
    chunk(o, "IEND", {});
    return o;
}

// ---- PNM (P5 / P6) and PFM ----
struct Header {
    std::string magic;
    int w, h;
    double maxval;
    size_t offset;
};

Header pnm_header(const std::vector<unsigned char> &d, int fields)
{
    Header hd{};
    std::string tok;
    std::vector<std::string> toks;
    size_t p = 0;
    while (p < d.size() && (int)toks.size() < fields) {
        const char ch = (char)d[p];
        if (ch == '#' && tok.empty()) {
            while (p < d.size() && d[p] != '\n') ++p;
        } else if (std::isspace((unsigned char)ch)) {
            if (!tok.empty()) toks.push_back(tok), tok.clear();
        } else {
            tok += ch;
        }
        ++p;
    }
    if ((int)toks.size() < fields) die("truncated image header");
    hd.magic = toks[0];
    auto side = [](const std::string &t) {
        // digits only, at most kMaxSide: no sign, no overflow in w * h
        if (t.empty() || t.size() > 6 || t.find_first_not_of("0123456789") != std::string::npos)
            die("bad image dimension '" + t + "'");
        const long v = std::stol(t);
        if (v < 1 || v > (long)kMaxSide) die("image dimension " + t + " out of range");
        return (int)v;
    };
    hd.w = side(toks[1]);
    hd.h = side(toks[2]);
    try {
        hd.maxval = std::stod(toks[3]);
    } catch (const std::exception &) {
        die("bad maxval '" + toks[3] + "'");
    }
    hd.offset = p;  // one whitespace byte after the last field was consumed
    return hd;
}

Image read_pnm(const std::vector<unsigned char> &d)
{
    const Header hd = pnm_header(d, 4);
    if (hd.magic != "P5" && hd.magic != "P6") die("unsupported PNM type " + hd.magic);
    const int spp = hd.magic == "P6" ? 3 : 1, bytes = hd.maxval > 255 ? 2 : 1;
    if (hd.maxval != 255 && hd.maxval != 65535) die("PNM maxval must be 255 or 65535");
    const size_t np = (size_t)hd.w * hd.h;
    if (d.size() < hd.offset + np * spp * bytes) die("truncated PNM data");
    Image im;
    im.w = hd.w;
    im.h = hd.h;
    im.depth = bytes * 8;
    im.grey_source = spp == 1;
    im.px.assign(3 * np, 0);
    const unsigned char *s = &d[hd.offset];
    for (size_t i = 0; i < np; ++i)
        for (int q = 0; q < 3; ++q) {
            const size_t o = (i * spp + (spp == 3 ? q : 0)) * bytes;
            im.px[q * np + i] = bytes == 1 ? s[o] : (s[o] << 8 | s[o + 1]);
        }
    return im;
}

std::vector<unsigned char> write_pnm(const Image &im, int depth)
{
    std::ostringstream h;
    h << "P6\n" << im.w << " " << im.h << "\n" << (depth == 8 ? 255 : 65535) << "\n";
    const std::string hs = h.str();
    std::vector<unsigned char> o(hs.begin(), hs.end());
    const size_t np = (size_t)im.w * im.h;
    for (size_t i = 0; i < np; ++i)
        for (int q = 0; q < 3; ++q) {
            const int v = im.px[q * np + i];
            if (depth == 16) o.push_back((unsigned char)(v >> 8));
            o.push_back((unsigned char)v);
        }
    return o;
}

Image read_pfm(const std::vector<unsigned char> &d)
{
    const Header hd = pnm_header(d, 4);
    if (hd.magic != "PF" && hd.magic != "Pf") die("not a PFM file");
    const int spp = hd.magic == "PF" ? 3 : 1;
    const bool little = hd.maxval < 0;
    const size_t np = (size_t)hd.w * hd.h;
    if (d.size() < hd.offset + np * spp * 4) die("truncated PFM data");
    Image im;
    im.w = hd.w;
    im.h = hd.h;
    im.depth = 32;
    im.grey_source = spp == 1;
    im.fx.assign(3 * np, 0.f);
    for (int y = 0; y < hd.h; ++y)       // PFM rows run bottom to top
        for (int x = 0; x < hd.w; ++x)
            for (int q = 0; q < 3; ++q) {
                const size_t o = hd.offset + (((size_t)(hd.h - 1 - y) * hd.w + x) * spp + (spp == 3 ? q : 0)) * 4;
                uint32_t u = little ? ((uint32_t)d[o] | (uint32_t)d[o + 1] << 8 | (uint32_t)d[o + 2] << 16 |
                                       (uint32_t)d[o + 3] << 24)
                                    : be32(&d[o]);
                float f;
                std::memcpy(&f, &u, 4);
                im.fx[q * np + (size_t)y * hd.w + x] = f;
            }
    return im;
}

std::vector<unsigned char> write_pfm(const Image &im)
{
    std::ostringstream h;
    h << "PF\n" << im.w << " " << im.h << "\n-1.0\n";
    const std::string hs = h.str();
    std::vector<unsigned char> o(hs.begin(), hs.end());
    const size_t np = (size_t)im.w * im.h;
    for (int y = im.h - 1; y >= 0; --y)
        for (int x = 0; x < im.w; ++x)
            for (int q = 0; q < 3; ++q) {
                uint32_t u;
                std::memcpy(&u, &im.fx[q * np + (size_t)y * im.w + x], 4);
                for (int s = 0; s < 32; s += 8) o.push_back((unsigned char)(u >> s));
            }
    return o;
}

std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

Image load(const std::string &path)
{
    const std::vector<unsigned char> d = read_file(path);
    if (d.size() >= 8 && d[0] == 137 && d[1] == 'P' && d[2] == 'N' && d[3] == 'G') return read_png(d);
    if (d.size() >= 2 && d[0] == 'P' && (d[1] == 'F' || d[1] == 'f')) return read_pfm(d);
    if (d.size() >= 2 && d[0] == 'P' && (d[1] == '5' || d[1] == '6')) return read_pnm(d);
    die("unknown image format: " + path);
}

// ---- arguments (BM3D_IO::arguments_process, include/BM3D.h:372-505) ----
struct Options {
    std::string filter = "--bm3d";
    std::string input, ref, tag = ".BM3D", format = ".png", profile = "fast", precision = "fp32";
    int device = 0;
    bool print_params = false;
    bool copy_only = false;   // read, convert and write without filtering (no GPU)
    isp_bm3d_params para{};
    isp_nlm_params nlm{};
};

int to_int(const std::string &opt, const std::vector<std::string> &a, size_t &i)
{
    if (++i >= a.size()) die("No argument specified for option " + opt + "!");
    return std::stoi(a[i]);
}

double to_double(const std::string &opt, const std::vector<std::string> &a, size_t &i)
{
    if (++i >= a.size()) die("No argument specified for option " + opt + "!");
    return std::stod(a[i]);
}

std::string to_str(const std::string &opt, const std::vector<std::string> &a, size_t &i)
{
    if (++i >= a.size()) die("No argument specified for option " + opt + "!");
    return a[i];
}

// Args::GetPara(int&, bool&) (include/Args.h:167-190): "true" / "false", any case
bool to_bool(const std::string &opt, const std::vector<std::string> &a, size_t &i)
{
    const std::string v = lower(to_str(opt, a, i));
    if (v == "true") return true;
    if (v == "false") return false;
    die("Invalid argument specified for option " + opt + ", must be \"true\" or \"false\"!");
}

Options parse(const std::vector<std::string> &a)
{
    Options o;
    // the profile first: it resets both steps' presets
    auto no_value = [](const std::string &s) { return s == "--print-params" || s == "--copy"; };
    for (size_t i = 0; i < a.size(); ++i) {
        if (a[i] == "-P" || a[i] == "--profile") o.profile = to_str(a[i], a, i);
        else if (!a[i].empty() && a[i][0] == '-' && !no_value(a[i])) ++i;
    }
    if (isp_bm3d_params_default(&o.para, o.profile.c_str()) != ISP_BM3D_OK) die("unknown profile " + o.profile);
    std::vector<double> sigma;
    bool th1 = false, th2 = false;
    isp_bm3d_step_params &b = o.para.basic, &f = o.para.final_;
    for (size_t i = 0; i < a.size(); ++i) {
        const std::string &s = a[i];
        if (s == "-P" || s == "--profile") ++i;
        else if (s == "-T" || s == "--tag") o.tag = to_str(s, a, i);
        else if (s == "-F" || s == "--format") o.format = to_str(s, a, i);
        else if (s == "--ref") o.ref = to_str(s, a, i);
        else if (s == "-S" || s == "--sigma") sigma.push_back(to_double(s, a, i));
        else if (s == "-BS1" || s == "--BlockSize1") b.block_size = to_int(s, a, i);
        else if (s == "-BSP1" || s == "--BlockStep1") b.block_step = to_int(s, a, i);
        else if (s == "-GS1" || s == "--GroupSize1") b.group_size = to_int(s, a, i);
        else if (s == "-MR1" || s == "--BMrange1") b.bm_range = to_int(s, a, i);
        else if (s == "-MS1" || s == "--BMstep1") b.bm_step = to_int(s, a, i);
        else if (s == "-TH1" || s == "--thMSE1") b.thMSE = to_double(s, a, i), th1 = true;
        else if (s == "-LD" || s == "--lambda") b.lambda = to_double(s, a, i);
        else if (s == "-BS2" || s == "--BlockSize2") f.block_size = to_int(s, a, i);
        else if (s == "-BSP2" || s == "--BlockStep2") f.block_step = to_int(s, a, i);
        else if (s == "-GS2" || s == "--GroupSize2") f.group_size = to_int(s, a, i);
        else if (s == "-MR2" || s == "--BMrange2") f.bm_range = to_int(s, a, i);
        else if (s == "-MS2" || s == "--BMstep2") f.bm_step = to_int(s, a, i);
        else if (s == "-TH2" || s == "--thMSE2") f.thMSE = to_double(s, a, i), th2 = true;
        else if (s == "-D" || s == "--device") o.device = to_int(s, a, i);
        else if (s == "--precision") {
            o.precision = to_str(s, a, i);
            (void)precision_code(o.precision);  // unknown names end the run here
        }
        else if (s == "--print-params") o.print_params = true;
        else if (s == "--copy") o.copy_only = true;
        else if (!s.empty() && s[0] == '-') die("unknown option " + s);
        else o.input = s;
    }
    // the default sigma is the preset's; given values repeat the last to three planes
    if (!sigma.empty()) {
        while (sigma.size() < 3) sigma.push_back(sigma.back());
        for (int c = 0; c < 3; ++c) b.sigma[c] = f.sigma[c] = sigma[c];
    } else {
        for (int c = 0; c < 3; ++c) f.sigma[c] = b.sigma[c];
    }
    if (!th1) isp_bm3d_thmse_default(&b, 0);
    if (!th2) isp_bm3d_thmse_default(&f, 1);
    if (o.input.empty() && !o.print_params) die("no input image given");
    if (o.format.empty() || o.format[0] != '.') o.format = "." + o.format;
    return o;
}

// NLMeans_IO::arguments_process (include/NLMeans.h:76-150) and FilterIO's -T / -F
Options parse_nlm(const std::vector<std::string> &a)
{
    Options o;
    o.filter = "--nlm";
    o.tag = ".NLMeans";
    isp_nlm_params_default(&o.nlm, 1, 8.0);
    isp_nlm_params &p = o.nlm;
    bool strength_def = false, th_def = false;
    for (size_t i = 0; i < a.size(); ++i) {
        const std::string &s = a[i];
        if (s == "-T" || s == "--tag") o.tag = to_str(s, a, i);
        else if (s == "-F" || s == "--format") o.format = to_str(s, a, i);
        else if (s == "--ref") o.ref = to_str(s, a, i);
        else if (s == "-C" || s == "--correction") p.correction = to_bool(s, a, i);
        else if (s == "-S" || s == "--sigma") p.sigma = to_double(s, a, i);
        else if (s == "-H" || s == "--strength") p.strength = to_double(s, a, i), strength_def = true;
        else if (s == "-BS" || s == "--BlockSize") p.block_size = to_int(s, a, i);
        else if (s == "-BSP" || s == "--BlockStep") p.block_step = to_int(s, a, i);
        else if (s == "-GS" || s == "--GroupSize") p.group_size = to_int(s, a, i);
        else if (s == "-MR" || s == "--BMrange") p.bm_range = to_int(s, a, i);
        else if (s == "-MS" || s == "--BMstep") p.bm_step = to_int(s, a, i);
        else if (s == "-TH" || s == "--thMSE") p.thMSE = to_double(s, a, i), th_def = true;
        else if (s == "-D" || s == "--device") o.device = to_int(s, a, i);
        else if (s == "--precision") {
            o.precision = to_str(s, a, i);
            (void)precision_code(o.precision);  // unknown names end the run here
        }
        else if (s == "--print-params") o.print_params = true;
        else if (s == "--copy") o.copy_only = true;
        else if (!s.empty() && s[0] == '-') die("unknown option " + s);
        else o.input = s;
    }
    // derived after parsing, from the final sigma and correction (:148-149)
    if (!strength_def) p.strength = p.correction ? p.sigma * 5 : p.sigma * 1.5;
    if (!th_def) p.thMSE = p.correction ? p.sigma * 50 : p.sigma * 25;
    if (o.input.empty() && !o.print_params) die("no input image given");
    if (o.format.empty() || o.format[0] != '.') o.format = "." + o.format;
    return o;
}

std::string output_path(const Options &o)
{
    const size_t slash = o.input.find_last_of("/\\");
    const std::string dir = slash == std::string::npos ? "" : o.input.substr(0, slash + 1);
    std::string name = slash == std::string::npos ? o.input : o.input.substr(slash + 1);
    const size_t dot = name.find_last_of('.');
    if (dot != std::string::npos) name = name.substr(0, dot);
    return dir + name + o.tag + o.format;
}

void print_params(const Options &o)
{
    if (o.filter != "--bm3d") {
        const isp_nlm_params &p = o.nlm;
        std::cout << "{\"filter\": \"nlm\", \"correction\": " << (p.correction ? "true" : "false")
                  << ", \"sigma\": " << p.sigma << ", \"strength\": " << p.strength << ", \"BlockSize\": "
                  << p.block_size << ", \"BlockStep\": " << p.block_step << ", \"GroupSize\": " << p.group_size
                  << ", \"BMrange\": " << p.bm_range << ", \"BMstep\": " << p.bm_step << ", \"thMSE\": " << p.thMSE
                  << ", \"input\": \"" << o.input << "\", \"ref\": \"" << o.ref << "\", \"output\": \""
                  << (o.input.empty() ? "" : output_path(o)) << "\", \"precision\": \"" << o.precision << "\"}"
                  << std::endl;
        return;
    }
    auto step = [](const isp_bm3d_step_params &p) {
        std::ostringstream s;
        s << "{\"sigma\": [" << p.sigma[0] << ", " << p.sigma[1] << ", " << p.sigma[2] << "], \"BlockSize\": "
          << p.block_size << ", \"BlockStep\": " << p.block_step << ", \"GroupSize\": " << p.group_size
          << ", \"BMrange\": " << p.bm_range << ", \"BMstep\": " << p.bm_step << ", \"thMSE\": " << p.thMSE
          << ", \"lambda\": " << p.lambda << "}";
        return s.str();
    };
    std::cout << "{\"profile\": \"" << o.profile << "\", \"basic\": " << step(o.para.basic) << ", \"final\": "
              << step(o.para.final_) << ", \"input\": \"" << o.input << "\", \"ref\": \"" << o.ref
              << "\", \"output\": \"" << (o.input.empty() ? "" : output_path(o)) << "\", \"precision\": \""
              << o.precision << "\"}" << std::endl;
}

// out: 16-bit frame samples, or float planes
int write_out(const Options &o, Image &out)
{
    const std::string opath = output_path(o), ext = lower(o.format);
    if (out.depth == 32) {
        if (ext != ".pfm") die("float images are written as .pfm");
        write_file(opath, write_pfm(out));
    } else {
        // ImageWriter: RoundDiv(value, 257) to 8 bits
        for (auto &v : out.px) v = (v + 128) / 257;
        if (ext == ".png") write_file(opath, write_png(out, 8));
        else if (ext == ".ppm" || ext == ".pnm") write_file(opath, write_pnm(out, 8));
        else die("unsupported output format " + o.format + " (.png, .ppm, .pfm)");
    }
    std::cout << opath << std::endl;
    return 0;
}

int precision_code(const std::string &p)
{
    if (p == "fp32") return ISP_BM3D_PREC_FP32;
    if (p == "fp16") return ISP_BM3D_PREC_FP16;
    if (p == "bf16") return ISP_BM3D_PREC_BF16;
    if (p == "fp32fast") return ISP_BM3D_PREC_FP32_FAST;
    if (p == "fp32mfma") return ISP_BM3D_PREC_FP32_MFMA;
    die("unknown precision " + p);
}

// ImageReader (include/ImageIO.h): 8-bit samples * 257 in a 16-bit frame (Floor 0, Ceil 65535)
void to_frame16(Image &im)
{
    const int scale = im.depth == 8 ? 257 : 1;
    for (auto &v : im.px) v *= scale;
}

int run_bm3d(const Options &o, Image &src, Image &ref, bool has_ref)
{
    isp_bm3d_ctx *ctx = nullptr;
    const int prec = precision_code(o.precision);
    if (isp_bm3d_create(&ctx, &o.para, o.device) != ISP_BM3D_OK) die(isp_bm3d_last_error(nullptr));
    int rc = isp_bm3d_set_precision(ctx, prec);
    const size_t np = (size_t)src.w * src.h;
    Image out = src;
    if (!rc && src.depth == 32) {
        // float frame in [0, 1]: the Frame semantics on float planes
        float *d[3], *s[3], *r[3];
        for (int c = 0; c < 3; ++c) {
            d[c] = out.fx.data() + c * np;
            s[c] = src.fx.data() + c * np;
            r[c] = has_ref ? ref.fx.data() + c * np : nullptr;
        }
        rc = isp_bm3d_rgb_f32(ctx, d[0], d[1], d[2], s[0], s[1], s[2], r[0], r[1], r[2], src.w, src.h, (size_t)src.w,
                              ISP_BM3D_BOTH, ISP_BM3D_MEM_HOST, nullptr);
    } else if (!rc) {
        to_frame16(src);
        if (has_ref) to_frame16(ref);
        int32_t *d[3], *s[3], *r[3];
        for (int c = 0; c < 3; ++c) {
            d[c] = out.px.data() + c * np;
            s[c] = src.px.data() + c * np;
            r[c] = has_ref ? ref.px.data() + c * np : nullptr;
        }
        rc = isp_bm3d_rgb_i32(ctx, d[0], d[1], d[2], s[0], s[1], s[2], r[0], r[1], r[2], src.w, src.h,
                              (size_t)src.w, 0, 65535, ISP_BM3D_BOTH, ISP_BM3D_MEM_HOST, nullptr);
    }
    if (rc) {
        const std::string m = isp_bm3d_last_error(ctx);
        isp_bm3d_destroy(ctx);
        die(m);
    }
    isp_bm3d_destroy(ctx);
    return write_out(o, out);
}

// NLMeans_IO::process (include/NLMeans.h:153-166): NLMeans::process_Frame on
// the 16-bit frame, guided by --ref when given
int run_nlm(const Options &o, Image &src, Image &ref, bool has_ref)
{
    if (src.depth == 32) die("NL-means filters integer frames (PNG / PNM), not float images");
    isp_bm3d_params bp{};
    isp_bm3d_params_default(&bp, "fast");
    const int prec = precision_code(o.precision);
    isp_bm3d_ctx *ctx = nullptr;
    if (isp_bm3d_create(&ctx, &bp, o.device) != ISP_BM3D_OK) die(isp_bm3d_last_error(nullptr));
    int rc = isp_bm3d_set_precision(ctx, prec);
    to_frame16(src);
    if (has_ref) to_frame16(ref);
    const size_t np = (size_t)src.w * src.h;
    Image out = src;
    if (!rc) {
        int32_t *d[3], *s[3], *r[3];
        for (int c = 0; c < 3; ++c) {
            d[c] = out.px.data() + c * np;
            s[c] = src.px.data() + c * np;
            r[c] = has_ref ? ref.px.data() + c * np : nullptr;
        }
        rc = isp_nlm_rgb_i32(ctx, &o.nlm, d[0], d[1], d[2], s[0], s[1], s[2], r[0], r[1], r[2], src.w, src.h,
                             (size_t)src.w, 0, 65535, ISP_BM3D_MEM_HOST, nullptr);
    }
    if (rc) {
        const std::string m = isp_bm3d_last_error(ctx);
        isp_bm3d_destroy(ctx);
        die(m);
    }
    isp_bm3d_destroy(ctx);
    return write_out(o, out);
}

// FilterIO::operator() (include/Filter.h:300-340): read, filter, write
int run(const Options &o)
{
    Image src = load(o.input);
    Image ref;
    const bool has_ref = !o.ref.empty();
    if (has_ref) {
        ref = load(o.ref);
        if (ref.w != src.w || ref.h != src.h || (ref.depth == 32) != (src.depth == 32))
            die("the reference image differs from the input in size or type");
    }
    if (o.copy_only) {
        Image out = src;
        if (out.depth != 32) to_frame16(out);
        return write_out(o, out);
    }
    return o.filter == "--bm3d" ? run_bm3d(o, src, ref, has_ref) : run_nlm(o, src, ref, has_ref);
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        if (argc <= 2) {
            std::cout << "Not enough arguments specified.\n";
            return 0;
        }
        // Filtering (source/ISP_MW.cpp:201-289): the filter name, lower-cased
        const std::string filter = lower(argv[1]);
        const bool nlm = filter == "--nlm" || filter == "--nlmeans" || filter == "--nonlocalmeans";
        if (filter != "--bm3d" && !nlm) {
            std::cerr << "isp_mw: only --bm3d and --nlm are part of this build\n";
            return 1;
        }
        const std::vector<std::string> args(argv + 2, argv + argc);
        const Options o = nlm ? parse_nlm(args) : parse(args);
        if (o.print_params) {
            print_params(o);
            return 0;
        }
        return run(o);
    } catch (const std::exception &e) {
        std::cerr << "isp_mw: " << e.what() << std::endl;
        return 2;
    }
}
