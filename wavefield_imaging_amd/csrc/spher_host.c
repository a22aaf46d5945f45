/*
 * Polar / spherical back end of libawefd (host C).  Placeholder until the
 * spherical kernels land: every entry point reports that the kind is not
 * available yet.
 */
#include "awefd_model.h"

int awefd__spher_create(awefd_model *m)
{
    (void)m;
    return awefd__fail(AWEFD_EINVAL, "polar/spherical kernels are not built yet");
}

int awefd__spher_set_model(awefd_model *m, const float *vel)
{
    (void)m; (void)vel;
    return awefd__fail(AWEFD_EINVAL, "polar/spherical kernels are not built yet");
}

int awefd__spher_set_sources(awefd_model *m)
{
    (void)m;
    return awefd__fail(AWEFD_EINVAL, "polar/spherical kernels are not built yet");
}

int awefd__spher_loop(awefd_model *m, float *dtr, int jsnap, awefd_snap_fn cb, void *user)
{
    (void)m; (void)dtr; (void)jsnap; (void)cb; (void)user;
    return awefd__fail(AWEFD_EINVAL, "polar/spherical kernels are not built yet");
}

int awefd__spher_get_field(awefd_model *m, float *host)
{
    (void)m; (void)host;
    return awefd__fail(AWEFD_EINVAL, "polar/spherical kernels are not built yet");
}

void awefd__spher_free(awefd_model *m) { (void)m; }
