/* qnb_oracle.h -- CPU restatement of the reference's nonbonded path.
 * Test infrastructure only (see qnb_oracle.c); the engine never links it. */
#ifndef QNB_ORACLE_H
#define QNB_ORACLE_H
#include "../include/qnb.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_handle orc_handle;

orc_handle *orc_create(const qnb_system *s, const qnb_params *p);
void orc_destroy(orc_handle *h);
void orc_nbqqlist(orc_handle *h);
void orc_make_pair_lists(orc_handle *h, const double *x);
int orc_eval_terms(orc_handle *h, uint32_t mask, const double *x, double *d, double Tfree, int istep,
                   qnb_energies *E, qnb_qenergies *EQ);
int64_t orc_get_pairs(orc_handle *h, int which, int32_t *ab, int64_t cap);
int64_t orc_get_list_params(orc_handle *h, int which, double *par, int64_t cap);
void orc_get_lrf(orc_handle *h, double *lrf);
void orc_set_lrf(orc_handle *h, const double *lrf);
void orc_get_shell_state(orc_handle *h, double *st);
void orc_set_shell_state(orc_handle *h, const double *st);
void orc_get_theta(orc_handle *h, double *theta, int32_t *n_insh);
int orc_calculate_exc(orc_handle *h, const double *x, int narr, const uint8_t *masks, const int32_t *types,
                      const double *eq_in, double *out);

#ifdef __cplusplus
}
#endif
#endif
