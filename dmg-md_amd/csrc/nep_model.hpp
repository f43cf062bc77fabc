// nep_model.hpp -- host representation of a parsed nep.txt and its device
// parameter image.
//
// Grammar and parameter layout follow the reference loader
// NEP::NEP(const char*) (src/gpumd_compat/nep.cu:135-415) and the pointer
// layout of NEP::update_potential (nep.cu:464-496).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace dmgx {

constexpr int kMaxTypes = 16;   // per-type device tables
constexpr int kMaxL = 8;        // 3-body angular order supported (nep_utilities.cuh:50, 80 moments)
constexpr int kPadL = 4;        // 3-body order of the bounded instances models are padded to
constexpr int kMaxNA = 10;      // n_max_angular + 1
constexpr int kMaxNR = 16;      // n_max_radial + 1
constexpr int kMaxK = 17;       // basis_size + 1 (nep_utilities.cuh MAX_NUM_N)
constexpr int kMaxDim = 128;
constexpr int kMaxNeurons = 128;

struct NepModel {
  int version = 4;
  int num_types = 0;
  std::vector<std::string> symbols;
  int Z[kMaxTypes] = {};
  bool zbl_enabled = false, zbl_flexible = false, zbl_typewise = false;
  double zbl_rc_inner = 0.0, zbl_rc_outer = 0.0, zbl_factor = 0.0;
  std::vector<double> zbl_para;  // 10 per unordered type pair (flexible ZBL)
  double rcR[kMaxTypes] = {}, rcA[kMaxTypes] = {};
  double rcR_max = 0.0, rcA_max = 0.0;
  int MN_R = 0, MN_A = 0;        // enlarged x1.25 (nep.cu:267-268)
  int nR = 0, nA = 0;            // n_max + 1
  int kR = 0, kA = 0;            // basis_size + 1
  int L_max = 0;
  int has222 = 0, has1111 = 0;
  int num_L = 0, dim = 0, H = 0;
  int num_para_ann = 0, num_para_desc = 0;
  std::vector<double> params;    // ann | descriptor (file order) | q_scaler

  // Parses `path`; throws std::invalid_argument with a message on any error.
  void load(const std::string& path);

  // Flat float image uploaded to the device; offsets in floats.
  struct Image {
    std::vector<float> data;
    int off_w0, off_b0, off_w1, off_w1x, off_b1, off_cR, off_cA, off_qs, off_zbl;
  };
  Image device_image() const;
};

// Closed-form angular constants: weights of s^2 in q_L (factor 2 of m>0
// folded in), integer coefficients of Z_{L,m}(z) = primitive part of
// d^m P_L/dz^m, and the q_222 / q_1111 weights (Wigner-3j couplings).
struct AngularConstants {
  float c3[(kMaxL + 1) * (kMaxL + 1) - 1];
  float zc[kMaxL + 1][kMaxL + 1][kMaxL + 1];
  float c4[5];
  float c5;
};
AngularConstants derive_angular_constants();

}  // namespace dmgx
