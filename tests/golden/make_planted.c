/* Planted probable primes for the full-Fermat workload (20,000-bit moduli, L = 2,048).
 *
 * Prime number i (i = first .. first + count - 1) is mpz_nextprime of a random 20,000-bit odd
 * number with the top bit set, from GMP's default generator seeded with 7000 + i, confirmed
 * by mpz_probab_prime_p with 40 rounds. Writes one hexadecimal prime per line.
 *
 * Build: gcc -O2 -I/opt/conda/include make_planted.c /usr/lib/x86_64-linux-gnu/libgmp.so.10
 * Run:   ./a.out FIRST COUNT BITS > part.txt
 */
#include <gmp.h>
#include <stdio.h>
#include <stdlib.h>

int main(int argc, char** argv) {
    long first = argc > 1 ? atol(argv[1]) : 0, count = argc > 2 ? atol(argv[2]) : 1;
    unsigned long bits = argc > 3 ? strtoul(argv[3], 0, 10) : 20000;
    mpz_t n, p;
    mpz_inits(n, p, NULL);
    gmp_randstate_t st;
    gmp_randinit_default(st);
    for (long i = first; i < first + count; i++) {
        gmp_randseed_ui(st, 7000 + (unsigned long)i);
        mpz_urandomb(n, st, bits);
        mpz_setbit(n, bits - 1);
        mpz_setbit(n, 0);
        mpz_nextprime(p, n);
        if (mpz_sizeinbase(p, 2) != bits || !mpz_probab_prime_p(p, 40)) {
            fprintf(stderr, "prime %ld: unexpected result\n", i);
            return 1;
        }
        gmp_printf("%Zx\n", p);
        fflush(stdout);
    }
    return 0;
}
