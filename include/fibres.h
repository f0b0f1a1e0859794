/* Reference fibre placement in liblippmannschwinger.so: the three functions
 * of the reference's libfibres.so, same names, arguments and results
 * (src/jaxmaterials/data/libfibres.cc:31-179; the reference compiles that
 * file with g++ at run time and binds it by ctypes in
 * src/jaxmaterials/data/distributions_fibres.py:121-170).  Host functions: no
 * GPU is involved and none needs to be present.
 *
 * The reference declares the state length as `unsigned int &`; the C form
 * below (a pointer) is the same ABI. */
#ifndef LS_FIBRES_H
#define LS_FIBRES_H

#ifdef __cplusplus
extern "C" {
#endif

/* Periodic distances from p to the n points q[2j], q[2j+1] in the box
 * L[0] x L[1]: sqrt of the squared distance minimised over the 3 x 3
 * periodic images, capped at 4 L[0] L[1] (libfibres.cc:31-46). */
void dist_periodic(const unsigned int n, const double p[2], const double *q,
                   const double L[2], double *dist);

/* Textual state of std::mt19937 seeded with `seed` into rng_state (at most
 * 7000 characters; the reference's caller passes a 10000-byte buffer); its
 * length into *rng_state_length (libfibres.cc:61-74). */
void initialise_rng(unsigned int seed, unsigned int *rng_state_length,
                    char *rng_state);

/* Yang Chen's random fibre movement (libfibres.cc:94-179): moves the
 * n_fibres fibres of radii r_fibres (fibre_locations[2n], in/out) in the
 * periodic box domain_size[2] until every fibre clears its neighbours by
 * eps_fibres, num_repeats sweeps at least; the random state is read from and
 * written back to rng_state / *rng_state_length.  Bit-identical results to
 * the reference's library for the same arguments and state.  Stops (with a
 * message on stderr) after 10^6 sweeps where the reference would not end. */
void move_fibres(const unsigned int n_fibres, const double *domain_size,
                 const double *r_fibres, double *fibre_locations,
                 const unsigned int num_repeats, const double eps_fibres,
                 const unsigned int it_max_ovlap,
                 unsigned int *rng_state_length, char *rng_state);

#ifdef __cplusplus
}
#endif

#endif /* LS_FIBRES_H */
