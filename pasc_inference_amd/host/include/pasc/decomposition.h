// Decomposition<HipVector>: the sizes of a problem (T time steps, R nodes, K clusters, xdim data
// channels) and the layout permutations of decomposition.cpp:246-356 on host arrays.  One GPU
// holds the whole problem, so the decomposition is the single-rank one (Tbegin = 0, Tend = T,
// every node local) and the permutations are plain index maps.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <vector>

#include "common.h"

namespace pascinference {

template <class VectorBase> class Decomposition;

template <> class Decomposition<HipVector> {
  public:
    Decomposition(int64_t T, int64_t R, int K, int xdim) : T_(T), R_(R), K_(K), xdim_(xdim) {
        if (T < 1 || R < 1 || K < 1 || xdim < 1) throw std::invalid_argument("Decomposition: sizes must be >= 1");
    }
    int64_t get_T() const { return T_; }
    int64_t get_R() const { return R_; }
    int get_K() const { return K_; }
    int get_xdim() const { return xdim_; }
    int64_t get_Tbegin() const { return 0; }
    int64_t get_Tend() const { return T_; }
    int64_t get_Tlocal() const { return T_; }

    // permute_TRblocksize (decomposition.cpp:265-356): component k of node r at time t moves
    // from (k*R + r)*T + t ("blockTR", orig) to (t*R + r)*bs + k ("TRblock", new); invert
    // copies new back to orig
    void permute_TRblocksize(std::vector<double> &orig, std::vector<double> &perm, int bs,
                             bool invert) const {
        const size_t n = (size_t)T_ * (size_t)R_ * (size_t)bs;
        if ((invert ? perm.size() : orig.size()) != n)
            throw std::invalid_argument("permute_TRblocksize: wrong vector length");
        (invert ? orig : perm).resize(n);
        for (int k = 0; k < bs; k++)
            for (int64_t r = 0; r < R_; r++)
                for (int64_t t = 0; t < T_; t++) {
                    const size_t a = ((size_t)k * R_ + r) * T_ + t;
                    const size_t b = ((size_t)t * R_ + r) * bs + k;
                    if (!invert) perm[b] = orig[a];
                    else orig[a] = perm[b];
                }
    }
    // data vectors (decomposition.cpp:246-252) and gamma vectors (decomposition.cpp:255-262)
    void permute_TRxdim(std::vector<double> &orig, std::vector<double> &perm, bool invert = false) const {
        permute_TRblocksize(orig, perm, xdim_, invert);
    }
    void permute_TRK(std::vector<double> &orig, std::vector<double> &perm, bool invert = false) const {
        permute_TRblocksize(orig, perm, K_, invert);
    }

  private:
    int64_t T_, R_;
    int K_, xdim_;
};

}  // namespace pascinference
