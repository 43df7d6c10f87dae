// Reference-style use of the C++ layer (include/isp_bm3d.hpp): denoise a
// small synthetic grey plane and RGB frame with BM3D and print a checksum.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "isp_bm3d.hpp"

int main()
{
    using namespace isp_bm3d;
    const int W = 64, H = 48;
    Plane_FL src(W, H);
    unsigned s = 12345;
    auto rnd = [&]() { s = s * 1103515245u + 12345u; return ((s >> 8) & 0xFFFF) / 65535.0f - 0.5f; };
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) src(y, x) = (x < W / 2 ? 0.3f : 0.7f) + 0.1f * rnd();
    try {
        BM3D_Para para("np");
        para.basic.sigma = {25, 25, 25};
        para.final.sigma = {25, 25, 25};
        para.thMSE_Default();
        BM3D filter(para);
        Plane_FL dst = filter(src);
        double sum = 0, err = 0;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                sum += dst(y, x);
                err += std::fabs(dst(y, x) - (x < W / 2 ? 0.3f : 0.7f));
            }
        Frame_FL f(W, H);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                f.R()(y, x) = src(y, x);
                f.G()(y, x) = 0.5f;
                f.B()(y, x) = 1.0f - src(y, x);
            }
        Frame_FL g = filter(f);
        std::printf("plane mean %.6f mean abs err %.6f frame R(0,0) %.6f\n", sum / (W * H), err / (W * H),
                    g.R()(0, 0));
        // integer frame, 8-bit full range: the same denoising through RangeConvert / OPP rounding
        Frame fi(W, H, 8, QuantRange::PC);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                fi.R()(y, x) = (int)std::lround(255.0 * std::fmin(1.0, std::fmax(0.0, src(y, x))));
                fi.G()(y, x) = 128;
                fi.B()(y, x) = 255 - fi.R()(y, x);
            }
        Frame go = filter(fi);
        double ierr = 0;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) ierr += std::fabs(go.R()(y, x) - 255.0 * (x < W / 2 ? 0.3 : 0.7));
        Plane pi(0, W, H, 10);  // 10-bit full-range luma plane
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) pi(y, x) = (int)std::lround(1023.0 * std::fmin(1.0, std::fmax(0.0, src(y, x))));
        Plane po = filter(pi);
        std::printf("int frame R mean abs err %.3f levels; int plane Ceil %d out(0,0) %d\n", ierr / (W * H), po.Ceil(),
                    po(0, 0));
        // NL-means of a grey integer frame (R = G = B; the frame above has a
        // constant luma, on which every distance is 0 and nothing matches)
        NLMeans nlm(NLMeans_Para(true, 20.0));
        Frame fg(W, H, 8, QuantRange::PC);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) fg.R()(y, x) = fg.G()(y, x) = fg.B()(y, x) = fi.R()(y, x);
        Frame gn = nlm(fg);
        double nerr = 0, noisy = 0;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                nerr += std::fabs(gn.R()(y, x) - 255.0 * (x < W / 2 ? 0.3 : 0.7));
                noisy += std::fabs(fi.R()(y, x) - 255.0 * (x < W / 2 ? 0.3 : 0.7));
            }
        Plane_FL pn = nlm(src);
        std::printf("nlm frame R mean abs err %.3f levels (noisy %.3f); nlm plane(0,0) %.6f\n", nerr / (W * H),
                    noisy / (W * H), pn(0, 0));
        BM3D_Final_Para fp("np");
        fp.BlockSize = 11;  // profile vn's final BlockSize: the generic kernels
        BM3D_Final fin(fp);
        fin(src, dst);
        std::printf("final step with BlockSize 11: out(0,0) %.6f\n", dst(0, 0));
        fp.BlockSize = 17;  // beyond every kernel: reported, not crashed
        BM3D_Final fin17(fp);
        try {
            fin17(src, dst);
            std::printf("unexpected success\n");
            return 1;
        } catch (const Error &e) {
            std::printf("expected error %d: %s\n", e.code, e.what());
        }
    } catch (const Error &e) {
        std::printf("error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
