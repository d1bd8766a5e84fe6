// Batches of ciphertext operations at mixed levels, executed as one fused set of launches
// per stage (fh_batch.hip).
//
// A batch collects EvalRotate, EvalMult (+ relinearization, optionally + Rescale), EvalMult by a
// plaintext (optionally + Rescale), EvalAdd (ciphertext or plaintext) and level-reduction items
// whose ciphertexts may sit at different levels. finalize() lays out one scratch arena,
// builds device job tables for every stage and uploads them once; run() then issues a fixed
// sequence of launches (tensor products, one INTT over every key-switch input limb, one ModUp
// conversion, one NTT over every ModUp target limb, one key inner product in which items that
// share a key read it once, one ModDown with the relinearization add fused into its last NTT,
// one Rescale) independent of the number of items, optionally captured once into a HIP graph
// and replayed. Every item's result is bit-identical to the single-operation entry points.
//
// The reference batches only same-level rotations sharing one key, in groups of 32
// (api/engine/cuda/CudaEngine.cpp:56-66, 656-702; src/CKKS/Ciphertext.cpp:1519-1599), and caches
// one CUDA graph per (level, moddown, slot) (src/CKKS/LimbPartitionMGPU.cu:669-892).
#pragma once
#include <memory>
#include <vector>

#include "fh_ops.hpp"

namespace fh {

struct BatchPlan;

class Batch {
 public:
  explicit Batch(Context& c);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  // ct, out: 2 x (l+1) limbs. key: device key ([dnum][2][L1+K][N], Montgomery form).
  void add_rotate(const uint64_t* ct, int l, const uint64_t* key, uint32_t galois, uint64_t* out);
  // EvalMult of a and b at level l (tensor + relinearization); with rescale the result is
  // rescaled to level l-1 (out: 2 x l limbs), else out has 2 x (l+1) limbs. towers_a / towers_b:
  // limbs per component of a / b when they sit above level l (0: l+1); their first l+1 limbs are read.
  void add_mult(const uint64_t* a, const uint64_t* b, int l, const uint64_t* relin_key, bool rescale, uint64_t* out, int towers_a = 0,
                int towers_b = 0);
  // out = a + b over ncomp components of l+1 limbs
  void add_add(const uint64_t* a, const uint64_t* b, int l, int ncomp, uint64_t* out);
  // ciphertext a (2 x (l+1) limbs) times plaintext pt (l+1 limbs), optionally rescaled
  void add_mult_plain(const uint64_t* a, const uint64_t* pt, int l, bool rescale, uint64_t* out);
  // (a0 + pt, a1) at level l
  void add_add_plain(const uint64_t* a, const uint64_t* pt, int l, uint64_t* out);
  // the first l_out+1 limbs of each of ncomp components of a (l_in+1 limbs each)
  void add_level_reduce(const uint64_t* a, int l_in, int l_out, int ncomp, uint64_t* out);
  // drop every item (keeping the scratch arena and table buffer for the next finalize)
  void reset();

  // Lay out scratch and upload the job tables (a copy ordered on the context stream after earlier
  // runs; the host does not wait for them). No items may be added afterwards.
  void finalize();
  // Issue the batch on st. graph: capture the launch sequence into a HIP graph on the first
  // call and replay it afterwards.
  void run(hipStream_t st, bool graph);
  size_t items() const;
  size_t scratch_bytes() const;

 private:
  Context& c_;
  std::unique_ptr<BatchPlan> p_;
};

}  // namespace fh
