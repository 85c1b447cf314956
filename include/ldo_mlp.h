/* ldo_mlp.h - the learned proposal's network: training and inference
 * weights on the device.
 *
 * The reference has no learned proposal; nothing here replaces a reference
 * interface.  The network scores where a staple end domain goes next to its
 * chain neighbour (the anchor), for the learned domain-regrowth move of the
 * sampler (ldo_learned_* in ldo.h):
 *
 *   input   a 7x7x7 box of lattice sites centred on the anchor, one-hot over
 *           13 channels per site (empty; scaffold x 6 orientations; staple x
 *           6 orientations), with the moving domain itself left out, plus a
 *           16-dim embedding of the moving domain's identity
 *           (|identity| - 1, i.e. the scaffold domain it complements)
 *   hidden  3 x 256, ReLU (BF16 MFMA, FP32 accumulate in training)
 *   output  the 6 sites next to the anchor x 6 orientations = 36 logits;
 *           every other site of the box is at distance > 1 from the anchor
 *           and never legal, so its logits are not formed
 *
 * Feature f of the one-hot input is site * 13 + channel with
 * site = (dz + 3) * 49 + (dy + 3) * 7 + (dx + 3) for an offset (dx, dy, dz)
 * from the anchor; channel 0 = empty, 1 + o = scaffold with orientation o,
 * 7 + o = staple with orientation o (o indexes +x,-x,+y,-y,+z,-z).  Output
 * k = 6 * n + o: neighbour n (the anchor plus unit vector n) and orientation
 * o.  A site holding a scaffold domain reads as scaffold; two domains of one
 * kind read as the smaller orientation index.
 *
 * Parameters are one flat FP32 vector in the order of ldo_mlp_layout;
 * matrices are row-major with the input index first (W[in][out]).
 */
#ifndef LDO_MLP_H
#define LDO_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDO_MLP_SITES 343
#define LDO_MLP_CHANNELS 13
#define LDO_MLP_FEATURES 4459
#define LDO_MLP_EMBED 16
#define LDO_MLP_HIDDEN 256
#define LDO_MLP_OUT 36

/* Element offsets of each tensor in the flat parameter vector. */
typedef struct ldo_mlp_layout {
    int64_t w1;  /* [LDO_MLP_FEATURES][LDO_MLP_HIDDEN] */
    int64_t emb; /* [n_ident][LDO_MLP_EMBED] */
    int64_t w1e; /* [LDO_MLP_EMBED][LDO_MLP_HIDDEN] */
    int64_t b1;  /* [LDO_MLP_HIDDEN] */
    int64_t w2;  /* [LDO_MLP_HIDDEN][LDO_MLP_HIDDEN] */
    int64_t b2;
    int64_t w3;
    int64_t b3;
    int64_t w4; /* [LDO_MLP_HIDDEN][LDO_MLP_OUT] */
    int64_t b4; /* [LDO_MLP_OUT] */
    int64_t total;
} ldo_mlp_layout;

/* Layout for a network with n_ident domain identities (>= 1). */
int ldo_mlp_layout_of(int32_t n_ident, ldo_mlp_layout* out);

/* A network on the current device.  params: a flat vector to start from, or
 * NULL for a seeded initialisation (uniform +-sqrt(6 / fan_in) weights, zero
 * biases, N(0, 1)-like embeddings).  Batches of up to max_batch samples. */
int ldo_mlp_create(int32_t n_ident, int32_t max_batch, const float* params, uint64_t seed,
                   void** out);
void ldo_mlp_destroy(void* mlp);
const char* ldo_mlp_last_error(void* mlp);
int ldo_mlp_get_params(void* mlp, float* out);
int ldo_mlp_set_params(void* mlp, const float* in);

/* One batch of training samples (host arrays):
 *   channels [batch][LDO_MLP_SITES]  uint8 channel of every site
 *   ident    [batch]                 identity index of the moving domain
 *   legal    [batch][LDO_MLP_OUT]    1 where the output may be proposed
 *   target   [batch]                 the output observed in the sample
 * Computes the mean cross-entropy of the masked softmax (loss) and its
 * gradient (kept on the device for ldo_mlp_adam_step; copied to grads when
 * that is not NULL, in the parameter layout).  The three hidden GEMMs and
 * their backward products run on bf16 MFMA tiles with FP32 accumulation;
 * the one-hot first layer is a gather. */
int ldo_mlp_loss_grad(void* mlp, int32_t batch, const uint8_t* channels, const int32_t* ident,
                      const uint8_t* legal, const int32_t* target, double* loss, float* grads);

/* Adam on the gradient of the last ldo_mlp_loss_grad. */
int ldo_mlp_adam_step(void* mlp, float lr, float beta1, float beta2, float eps);

/* Masked log-probabilities [batch][LDO_MLP_OUT] (-INFINITY where illegal),
 * same arithmetic as training. */
int ldo_mlp_log_probs(void* mlp, int32_t batch, const uint8_t* channels, const int32_t* ident,
                      const uint8_t* legal, float* out);

#ifdef __cplusplus
}
#endif

#endif
