// ref_set_encoder_shim.cpp -- exports the reference's own colex codec
// (Treedepth/src/set_encoder/set_encoder.{hpp,cu}, compiled from where it
// lies under the reference tree) through a C ABI, so the restatement in
// oracle_core.cpp can be compared with it.  Test infrastructure only; built
// into oracle/_ref/ by oracle/Makefile when the reference tree is present.
#include <cstdint>

#include "src/set_encoder/set_encoder.hpp"

extern "C" {
uint64_t ref_nchoosek(uint64_t n, uint64_t k) {
  return td::set_encoder::NChooseK(n, k);
}
uint64_t ref_encode(const int32_t* s, int32_t k) {
  return td::set_encoder::Encode(s, static_cast<std::size_t>(k));
}
uint64_t ref_encode_exclude(const int32_t* s, int32_t k, int32_t ex) {
  return td::set_encoder::Encode(s, static_cast<std::size_t>(k),
                                 static_cast<std::size_t>(ex));
}
void ref_decode(int32_t n, int32_t k, uint64_t code, int32_t* out) {
  td::set_encoder::Decode(static_cast<std::size_t>(n),
                          static_cast<std::size_t>(k), code, out);
}
}
