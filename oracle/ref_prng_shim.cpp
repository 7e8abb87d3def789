// Exposes the reference's own host md5Rand (src/padenti/prng.hpp:107-195),
// compiled unchanged from the reference sources, through a C symbol so the
// oracle's MD5 restatement can be checked against it.  Test infrastructure.
#include <padenti/prng.hpp>

extern "C" void ref_md5rand(const unsigned int seed[4], unsigned int state[4])
{
  md5Rand(seed, state);
}
