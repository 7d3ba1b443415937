// ImageData<HipVector>: an image read from a PETSc binary Vec (imagedata.cpp:7-40), its pixels
// the R = width x height nodes of one time step on the 4-neighbour grid, and saveImage
// (imagedata.cpp:43-113).
#pragma once

#include "decomposition.h"
#include "tsdata.h"

namespace pascinference {
namespace data {

template <class VectorBase> class ImageData;

template <> class ImageData<HipVector> : public TSData<HipVector> {
  public:
    // The file holds the pixels in the original layout (blockTR: channel-major, row by row);
    // permute_TRxdim maps them to the node-major layout of the problem (an identity for one
    // channel at T = 1).
    ImageData(pasc_ctx *ctx, const Decomposition<HipVector> &decomposition,
              const std::string &filename, int width, int height, int precision = PASC_FP64)
        : TSData<HipVector>(ctx, load(decomposition, filename, width, height),
                            decomposition.get_K(), decomposition.get_xdim(), precision, nullptr,
                            1.0, PASC_FEM_HAT, (int64_t)width * height, width),
          decomposition_(decomposition), width_(width), height_(height) {}

    int get_width() const { return width_; }
    int get_height() const { return height_; }

    // saveImage (imagedata.cpp:43-113): <prefix>_original.bin (when save_original) and the
    // recovered image sum_k theta_k gamma_k as <prefix>_recovered.bin, both permuted back by
    // permute_TRxdim, and gamma permuted back by permute_TRK (k-major) as <prefix>_gamma.bin
    void saveImage(const std::string &prefix, bool save_original) const {
        const Decomposition<HipVector> &dec = decomposition_;
        std::vector<double> data = x_, out;
        if (save_original) {
            dec.permute_TRxdim(out, data, true);
            petscio::write_vec(prefix + "_original.bin", out);
        }
        std::vector<double> g = get_gammavector(), gout;
        dec.permute_TRK(gout, g, true);
        petscio::write_vec(prefix + "_gamma.bin", gout);
        const std::vector<double> th = get_thetavector();
        const size_t n = (size_t)(T_ * R_);
        std::vector<double> rec(n * xdim_, 0.0);
        for (size_t i = 0; i < n; i++)
            for (int k = 0; k < K_; k++)
                for (int c = 0; c < xdim_; c++)
                    rec[i * xdim_ + c] += th[(size_t)k * xdim_ + c] * g[i * K_ + k];
        dec.permute_TRxdim(out, rec, true);
        petscio::write_vec(prefix + "_recovered.bin", out);
    }

  private:
    static std::vector<double> load(const Decomposition<HipVector> &decomposition,
                                    const std::string &filename, int width, int height) {
        std::vector<double> raw = petscio::read_vec(filename), x;
        if (decomposition.get_T() != 1 || decomposition.get_R() != (int64_t)width * height)
            throw std::invalid_argument("ImageData: the decomposition is not T = 1 of width x height nodes");
        if ((int64_t)raw.size() != (int64_t)width * height * decomposition.get_xdim())
            throw std::runtime_error("ImageData: " + filename + " does not hold width x height x xdim values");
        decomposition.permute_TRxdim(raw, x);
        return x;
    }

    Decomposition<HipVector> decomposition_;
    int width_, height_;
};

}  // namespace data
}  // namespace pascinference
