// isp_bm3d.hpp - header-only C++ layer over the C ABI (isp_bm3d.h) with the
// class names of ISP_MW's BM3D interface, so that reference-style code reads
// the same:
//
//     isp_bm3d::BM3D_Para para("np");             // BM3D_Para, include/BM3D.h:154-168
//     para.basic.sigma = {25, 25, 25};
//     para.thMSE_Default();
//     isp_bm3d::BM3D filter(para);                // BM3D, include/BM3D.h:337-357
//     isp_bm3d::Plane_FL dst = filter(src);       // FilterIF2::operator(), include/Filter.h:211-216
//
// Plane_FL holds float samples (luma range [0,1]) with stride == width, like
// the reference's Plane_FL (include/Image_Type.h:186-290).  Plane holds int32
// samples with BitDepth and Floor / Neutral / Ceil (include/Image_Type.h:51-183)
// and Frame three such RGB planes (include/Image_Type.h:294-403); BM3D on them
// converts with RangeConvert / the OPP matrices and rounds back exactly as the
// reference's process_Plane / process_Frame do.  Frame_FL holds three float RGB
// planes in [0,1] with the same process_Frame semantics.
// NLMeans_Para / NLMeans (include/NLMeans.h) run NL-means on the same
// matching kernel, on Plane_FL, Plane and Frame.
// Errors throw isp_bm3d::Error (the reference prints and exits instead).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "isp_bm3d.h"
#include "isp_nlm.h"

namespace isp_bm3d {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const isp_bm3d_ctx *ctx = nullptr)
{
    if (rc < 0) throw Error(rc, isp_bm3d_last_error(ctx));
}

// BM3D_Para_Base (include/BM3D.h:15-45) and its two step presets.
struct BM3D_Para_Base {
    std::string profile;
    std::vector<double> sigma;
    int BlockSize, BlockStep, GroupSize, BMrange, BMstep;
    double thMSE, lambda;
    bool final_step;

    BM3D_Para_Base(const std::string &prof, bool fin) : profile(prof), final_step(fin)
    {
        isp_bm3d_step_params p;
        check(isp_bm3d_step_params_default(&p, prof.c_str(), fin ? 1 : 0));
        from_c(p);
    }
    void thMSE_Default()
    {
        isp_bm3d_step_params p = to_c();
        check(isp_bm3d_thmse_default(&p, final_step ? 1 : 0));
        thMSE = p.thMSE;
    }
    isp_bm3d_step_params to_c() const
    {
        isp_bm3d_step_params p{};
        profile.copy(p.profile, sizeof(p.profile) - 1);
        for (int c = 0; c < 3; ++c)
            p.sigma[c] = sigma.empty() ? 0.0 : sigma[c < (int)sigma.size() ? c : sigma.size() - 1];
        p.block_size = BlockSize;
        p.block_step = BlockStep;
        p.group_size = GroupSize;
        p.bm_range = BMrange;
        p.bm_step = BMstep;
        p.thMSE = thMSE;
        p.lambda = lambda;
        return p;
    }
    void from_c(const isp_bm3d_step_params &p)
    {
        sigma.assign(p.sigma, p.sigma + 3);
        BlockSize = p.block_size;
        BlockStep = p.block_step;
        GroupSize = p.group_size;
        BMrange = p.bm_range;
        BMstep = p.bm_step;
        thMSE = p.thMSE;
        lambda = p.lambda;
    }
};

struct BM3D_Basic_Para : BM3D_Para_Base {
    explicit BM3D_Basic_Para(const std::string &profile = "fast") : BM3D_Para_Base(profile, false) {}
};

struct BM3D_Final_Para : BM3D_Para_Base {
    explicit BM3D_Final_Para(const std::string &profile = "fast") : BM3D_Para_Base(profile, true) {}
};

struct BM3D_Para {
    BM3D_Basic_Para basic;
    BM3D_Final_Para final;
    explicit BM3D_Para(const std::string &profile = "fast") : basic(profile), final(profile) {}
    void thMSE_Default()
    {
        basic.thMSE_Default();
        final.thMSE_Default();
    }
};

class Plane_FL {
public:
    Plane_FL() = default;
    Plane_FL(int width, int height, float value = 0.f)
        : w_(width), h_(height), data_((size_t)width * height, value) {}
    int Width() const { return w_; }
    int Height() const { return h_; }
    int Stride() const { return w_; }
    float *data() { return data_.data(); }
    const float *data() const { return data_.data(); }
    float &operator()(int y, int x) { return data_[(size_t)y * w_ + x]; }
    float operator()(int y, int x) const { return data_[(size_t)y * w_ + x]; }

private:
    int w_ = 0, h_ = 0;
    std::vector<float> data_;
};

// QuantRange and the quantisation of a bit depth (include/Image_Type.h:39-44,
// include/Image_Type.hpp:94-109).
enum class QuantRange { TV, PC };

inline void Quantize_Value(int &Floor, int &Neutral, int &Ceil, int BitDepth, QuantRange qr, bool Chroma)
{
    if (BitDepth < 8 || BitDepth > 30) throw Error(ISP_BM3D_E_ARG, "bit depth must be 8..30");
    if (qr == QuantRange::PC) {
        Floor = 0;
        Ceil = (1 << BitDepth) - 1;
    } else {
        Floor = 16 << (BitDepth - 8);
        Ceil = (Chroma ? 240 : 235) << (BitDepth - 8);
    }
    Neutral = Chroma ? 1 << (BitDepth - 1) : Floor;
}

class Plane {
public:
    Plane() = default;
    // Plane(Value, Width, Height, BitDepth): full-range luma quantisation
    explicit Plane(int Value, int width = 1920, int height = 1080, int BitDepth = 16)
        : w_(width), h_(height), bd_(BitDepth), data_((size_t)width * height, Value)
    {
        Quantize_Value(floor_, neutral_, ceil_, BitDepth, QuantRange::PC, false);
    }
    Plane(int Value, int width, int height, int BitDepth, int Floor, int Neutral, int Ceil)
        : w_(width), h_(height), bd_(BitDepth), floor_(Floor), neutral_(Neutral), ceil_(Ceil),
          data_((size_t)width * height, Value) {}
    int Width() const { return w_; }
    int Height() const { return h_; }
    int Stride() const { return w_; }
    int BitDepth() const { return bd_; }
    int Floor() const { return floor_; }
    int Neutral() const { return neutral_; }
    int Ceil() const { return ceil_; }
    int ValueRange() const { return ceil_ - floor_; }
    bool isChroma() const { return floor_ < neutral_; }
    bool isPCChroma() const { return floor_ < neutral_ && (floor_ + ceil_) % 2 == 1; }
    int32_t *data() { return data_.data(); }
    const int32_t *data() const { return data_.data(); }
    int32_t &operator()(int y, int x) { return data_[(size_t)y * w_ + x]; }
    int32_t operator()(int y, int x) const { return data_[(size_t)y * w_ + x]; }
    // a plane of the same size and quantisation (what the reference's operator() allocates)
    Plane like() const { return Plane(0, w_, h_, bd_, floor_, neutral_, ceil_); }

private:
    int w_ = 0, h_ = 0, bd_ = 16, floor_ = 0, neutral_ = 0, ceil_ = 65535;
    std::vector<int32_t> data_;
};

// RGB frame of int32 planes (R, G, B are luma-type: Floor == Neutral).
class Frame {
public:
    Frame() = default;
    Frame(int width, int height, int BitDepth = 16, QuantRange qr = QuantRange::PC)
    {
        int fl, ne, ce;
        Quantize_Value(fl, ne, ce, BitDepth, qr, false);
        for (auto &p : p_) p = Plane(fl, width, height, BitDepth, fl, ne, ce);
    }
    Plane &R() { return p_[0]; }
    Plane &G() { return p_[1]; }
    Plane &B() { return p_[2]; }
    const Plane &R() const { return p_[0]; }
    const Plane &G() const { return p_[1]; }
    const Plane &B() const { return p_[2]; }
    int Width() const { return p_[0].Width(); }
    int Height() const { return p_[0].Height(); }
    int BitDepth() const { return p_[0].BitDepth(); }
    Frame like() const
    {
        Frame f;
        for (int c = 0; c < 3; ++c) f.p_[c] = p_[c].like();
        return f;
    }

private:
    Plane p_[3];
};

class Frame_FL {
public:
    Frame_FL() = default;
    Frame_FL(int width, int height) : p_{Plane_FL(width, height), Plane_FL(width, height), Plane_FL(width, height)} {}
    Plane_FL &R() { return p_[0]; }
    Plane_FL &G() { return p_[1]; }
    Plane_FL &B() { return p_[2]; }
    const Plane_FL &R() const { return p_[0]; }
    const Plane_FL &G() const { return p_[1]; }
    const Plane_FL &B() const { return p_[2]; }
    int Width() const { return p_[0].Width(); }
    int Height() const { return p_[0].Height(); }

private:
    Plane_FL p_[3];
};

// FilterIF2-style wrapper (include/Filter.h:138-217) around one context.
class FilterBase {
public:
    FilterBase(const FilterBase &) = delete;
    FilterBase &operator=(const FilterBase &) = delete;
    ~FilterBase() { isp_bm3d_destroy(ctx_); }

    Plane_FL &process(Plane_FL &dst, const Plane_FL &src, const Plane_FL &ref)
    {
        size_check(dst, src);
        size_check(ref, src);
        check(isp_bm3d_plane_f32(ctx_, dst.data(), src.data(), ref.data(), src.Width(), src.Height(), src.Stride(),
                                 mode_, ISP_BM3D_MEM_HOST, nullptr),
              ctx_);
        return dst;
    }
    Plane_FL &process(Plane_FL &dst, const Plane_FL &src) { return process(dst, src, src); }
    Frame_FL &process(Frame_FL &dst, const Frame_FL &src, const Frame_FL &ref)
    {
        size_check(dst.R(), src.R());
        size_check(ref.R(), src.R());
        check(isp_bm3d_rgb_f32(ctx_, dst.R().data(), dst.G().data(), dst.B().data(), src.R().data(), src.G().data(),
                               src.B().data(), ref.R().data(), ref.G().data(), ref.B().data(), src.Width(),
                               src.Height(), src.R().Stride(), mode_, ISP_BM3D_MEM_HOST, nullptr),
              ctx_);
        return dst;
    }
    Frame_FL &process(Frame_FL &dst, const Frame_FL &src) { return process(dst, src, src); }

    // integer Plane: RangeConvert in and out (source/BM3D.cpp:248-279)
    Plane &process(Plane &dst, const Plane &src, const Plane &ref)
    {
        size_check(dst, src);
        size_check(ref, src);
        check(isp_bm3d_plane_i32(ctx_, dst.data(), src.data(), ref.data(), src.Width(), src.Height(), src.Stride(),
                                 src.Floor(), src.Neutral(), src.Ceil(), mode_, ISP_BM3D_MEM_HOST, nullptr),
              ctx_);
        return dst;
    }
    Plane &process(Plane &dst, const Plane &src) { return process(dst, src, src); }
    // integer RGB Frame: OPP in, OPP out with +0.5 and clip (source/BM3D.cpp:282-313)
    Frame &process(Frame &dst, const Frame &src, const Frame &ref)
    {
        size_check(dst.R(), src.R());
        size_check(ref.R(), src.R());
        const Plane &q = src.R();
        if (q.Floor() != q.Neutral()) throw Error(ISP_BM3D_E_ARG, "an RGB frame has luma-type planes");
        check(isp_bm3d_rgb_i32(ctx_, dst.R().data(), dst.G().data(), dst.B().data(), src.R().data(), src.G().data(),
                               src.B().data(), ref.R().data(), ref.G().data(), ref.B().data(), src.Width(),
                               src.Height(), q.Stride(), q.Floor(), q.Ceil(), mode_, ISP_BM3D_MEM_HOST, nullptr),
              ctx_);
        return dst;
    }
    Frame &process(Frame &dst, const Frame &src) { return process(dst, src, src); }

    template <typename T>
    T operator()(const T &src)
    {
        T dst = make_like(src);
        process(dst, src);
        return dst;
    }
    template <typename T>
    T operator()(const T &src, const T &ref)
    {
        T dst = make_like(src);
        process(dst, src, ref);
        return dst;
    }
    isp_bm3d_ctx *context() { return ctx_; }

protected:
    FilterBase(const isp_bm3d_params &p, int mode, int device) : mode_(mode)
    {
        check(isp_bm3d_create(&ctx_, &p, device));
    }
    template <typename P>
    static void size_check(const P &a, const P &b)
    {
        if (a.Width() != b.Width() || a.Height() != b.Height())
            throw Error(ISP_BM3D_E_ARG, "planes of different sizes");
    }
    static Plane_FL make_like(const Plane_FL &p) { return Plane_FL(p.Width(), p.Height()); }
    static Frame_FL make_like(const Frame_FL &f) { return Frame_FL(f.Width(), f.Height()); }
    static Plane make_like(const Plane &p) { return p.like(); }
    static Frame make_like(const Frame &f) { return f.like(); }
    isp_bm3d_ctx *ctx_ = nullptr;
    int mode_;
};

// BM3D_Basic (include/BM3D.h:287-309)
class BM3D_Basic : public FilterBase {
public:
    explicit BM3D_Basic(const BM3D_Basic_Para &para = BM3D_Basic_Para(), int device = 0)
        : FilterBase(make(para), ISP_BM3D_BASIC, device) {}

private:
    static isp_bm3d_params make(const BM3D_Basic_Para &para)
    {
        isp_bm3d_params p;
        check(isp_bm3d_params_default(&p, para.profile.c_str()));
        p.basic = para.to_c();
        return p;
    }
};

// BM3D_Final (include/BM3D.h:312-334)
class BM3D_Final : public FilterBase {
public:
    explicit BM3D_Final(const BM3D_Final_Para &para = BM3D_Final_Para(), int device = 0)
        : FilterBase(make(para), ISP_BM3D_FINAL, device) {}

private:
    static isp_bm3d_params make(const BM3D_Final_Para &para)
    {
        isp_bm3d_params p;
        check(isp_bm3d_params_default(&p, para.profile.c_str()));
        p.final_ = para.to_c();
        return p;
    }
};

// BM3D (include/BM3D.h:337-357)
class BM3D : public FilterBase {
public:
    explicit BM3D(const BM3D_Para &para = BM3D_Para(), int device = 0)
        : FilterBase(make(para), ISP_BM3D_BOTH, device) {}

private:
    static isp_bm3d_params make(const BM3D_Para &para)
    {
        isp_bm3d_params p;
        p.basic = para.basic.to_c();
        p.final_ = para.final.to_c();
        return p;
    }
};

// NLMeans_Para (include/NLMeans.h:11-22): strength and thMSE follow sigma and
// correction unless set; Derive() recomputes them as NLMeans_IO does (:148-149).
struct NLMeans_Para {
    bool correction = true;
    double sigma = 8.0;
    double strength = 40.0;
    int BlockSize = 8, BlockStep = 5, GroupSize = 16, BMrange = 24, BMstep = 3;
    double thMSE = 400.0;

    NLMeans_Para() = default;
    NLMeans_Para(bool corr, double s) : correction(corr), sigma(s) { Derive(); }
    void Derive()
    {
        strength = correction ? sigma * 5 : sigma * 1.5;
        thMSE = correction ? sigma * 50 : sigma * 25;
    }
    isp_nlm_params to_c() const
    {
        isp_nlm_params p{};
        p.correction = correction;
        p.sigma = sigma;
        p.strength = strength;
        p.block_size = BlockSize;
        p.block_step = BlockStep;
        p.group_size = GroupSize;
        p.bm_range = BMrange;
        p.bm_step = BMstep;
        p.thMSE = thMSE;
        return p;
    }
};

// NLMeans (include/NLMeans.h:25-59): FilterIF2-style process / operator() on
// Plane_FL (process_Plane_FL), Plane (process_Plane) and RGB Frame
// (process_Frame, matching on the OPP luma of ref).
class NLMeans {
public:
    explicit NLMeans(const NLMeans_Para &para = NLMeans_Para(), int device = 0) : para_(para.to_c())
    {
        isp_bm3d_params p;
        check(isp_bm3d_params_default(&p, "fast"));
        check(isp_bm3d_create(&ctx_, &p, device));
    }
    NLMeans(const NLMeans &) = delete;
    NLMeans &operator=(const NLMeans &) = delete;
    ~NLMeans() { isp_bm3d_destroy(ctx_); }

    Plane_FL &process(Plane_FL &dst, const Plane_FL &src, const Plane_FL &ref)
    {
        same(dst, src);
        same(ref, src);
        check(isp_nlm_plane_f32(ctx_, &para_, dst.data(), src.data(), ref.data(), src.Width(), src.Height(),
                                src.Stride(), ISP_BM3D_MEM_HOST, nullptr),
              ctx_);
        return dst;
    }
    Plane &process(Plane &dst, const Plane &src, const Plane &ref)
    {
        same(dst, src);
        same(ref, src);
        check(isp_nlm_plane_i32(ctx_, &para_, dst.data(), src.data(), ref.data(), src.Width(), src.Height(),
                                src.Stride(), src.Floor(), src.Neutral(), src.Ceil(), ISP_BM3D_MEM_HOST, nullptr),
              ctx_);
        return dst;
    }
    Frame &process(Frame &dst, const Frame &src, const Frame &ref)
    {
        same(dst.R(), src.R());
        same(ref.R(), src.R());
        const Plane &q = src.R();
        check(isp_nlm_rgb_i32(ctx_, &para_, dst.R().data(), dst.G().data(), dst.B().data(), src.R().data(),
                              src.G().data(), src.B().data(), ref.R().data(), ref.G().data(), ref.B().data(),
                              src.Width(), src.Height(), q.Stride(), q.Floor(), q.Ceil(), ISP_BM3D_MEM_HOST, nullptr),
              ctx_);
        return dst;
    }
    template <typename T>
    T &process(T &dst, const T &src) { return process(dst, src, src); }
    Plane_FL operator()(const Plane_FL &src) { return run(Plane_FL(src.Width(), src.Height()), src, src); }
    Plane_FL operator()(const Plane_FL &src, const Plane_FL &ref)
    {
        return run(Plane_FL(src.Width(), src.Height()), src, ref);
    }
    Plane operator()(const Plane &src) { return run(src.like(), src, src); }
    Plane operator()(const Plane &src, const Plane &ref) { return run(src.like(), src, ref); }
    Frame operator()(const Frame &src) { return run(src.like(), src, src); }
    Frame operator()(const Frame &src, const Frame &ref) { return run(src.like(), src, ref); }
    isp_bm3d_ctx *context() { return ctx_; }

private:
    template <typename T>
    T run(T dst, const T &src, const T &ref)
    {
        process(dst, src, ref);
        return dst;
    }
    template <typename P>
    static void same(const P &a, const P &b)
    {
        if (a.Width() != b.Width() || a.Height() != b.Height())
            throw Error(ISP_BM3D_E_ARG, "planes of different sizes");
    }
    isp_nlm_params para_;
    isp_bm3d_ctx *ctx_ = nullptr;
};

}  // namespace isp_bm3d
