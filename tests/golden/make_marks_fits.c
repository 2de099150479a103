/* Writes tests/golden/marks_cfitsio.fits with cfitsio, making the same calls in
 * the same order as the reference's create_marks_fits_file
 * (src/input_formats.f90:605-694): primary header keys, then one BINTABLE per
 * field with columns x, y, w, xi_E_MARKS, xi_B_MARKS, K_MARKS (FP32 here) and a
 * LENSFILE long string.  Small deterministic contents; a fixture for the
 * numpy reader in 2pcf_amd/marks_io.py.
 *
 *   gcc -O1 tests/golden/make_marks_fits.c -I/opt/conda/include -L/opt/conda/lib -lcfitsio \
 *       -Wl,-rpath,/opt/conda/lib -o /tmp/make_marks_fits && /tmp/make_marks_fits tests/golden/marks_cfitsio.fits
 *
 * With a second argument, reads that file instead and prints every value
 * (used to check files written by marks_io.py with cfitsio itself). */
#include <fitsio.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NBIN 4

static void check(int status)
{
    if (status) {
        fits_report_error(stderr, status);
        exit(1);
    }
}

static int write_file(const char *name)
{
    fitsfile *f;
    int status = 0;
    char fname[512];
    snprintf(fname, sizeof fname, "!%s", name);
    ffinit(&f, fname, &status);
    check(status);
    ffphps(f, 8, 0, NULL, &status);
    ffpkys(f, "SURVEY", "", "", &status);
    ffpkyj(f, "NFIELD", 2, "Number of fields", &status);
    ffpkyj(f, "NBIN", NBIN, "Number of angular bins", &status);
    ffpkyj(f, "BINTYPE", 1, "Bin spacing. lin = 0, log = 1, ihs = 2", &status);
    ffpkye(f, "THETAMIN", 1.0f, 7, "Innermost bin edge", &status);
    ffpkye(f, "THETAMAX", 16.0f, 7, "Outermost bin edge", &status);
    float ce[NBIN] = {1.5f, 3.0f, 6.0f, 12.0f}, ed[NBIN + 1] = {1.0f, 2.0f, 4.0f, 8.0f, 16.0f},
          th[NBIN] = {1.45f, 2.9f, 5.8f, 11.6f};
    ffpkne(f, "BINCE", 1, NBIN, ce, 7, (char **)(const char *[]){"Bin center&&"}, &status);
    ffpkne(f, "BINED", 1, NBIN + 1, ed, 7, (char **)(const char *[]){"Bin edge&&"}, &status);
    ffpkne(f, "BINTH", 1, NBIN, th, 7, (char **)(const char *[]){"Mean theta&&"}, &status);
    check(status);
    const int nobj[2] = {5, 7};
    for (int i = 0; i < 2; ++i) {
        const int n = nobj[i];
        char extname[16], form[16];
        snprintf(extname, sizeof extname, "FIELD%3d", i + 1);
        snprintf(form, sizeof form, "%dE", NBIN);
        char *ttype[1] = {"x"}, *tform[1] = {"E"}, *tunit[1] = {""};
        ffibin(f, n, 1, ttype, tform, tunit, extname, 0, &status);
        fficol(f, 2, "y", "E", &status);
        fficol(f, 3, "w", "E", &status);
        fficol(f, 4, "xi_E_MARKS", form, &status);
        fficol(f, 5, "xi_B_MARKS", form, &status);
        fficol(f, 6, "K_MARKS", form, &status);
        check(status);
        ffpkls(f, "LENSFILE",
               i == 0 ? "lenses/field_1.fits"
                      : "a/rather/long/directory/name/for/the/second/field/lens/catalogue/of/this/test/field_2.fits",
               "", &status);
        ffplsw(f, &status);
        double x[8], y[8], w[8], e[8 * NBIN], b[8 * NBIN], k[8 * NBIN];
        for (int j = 0; j < n; ++j) {
            x[j] = 1.5 * j + i;
            y[j] = 0.25 * j - i;
            w[j] = 1.0 + 0.125 * j;
            for (int q = 0; q < NBIN; ++q) {
                e[j * NBIN + q] = 0.0625 * (q + 1) * (j + 1) * (i + 1);
                b[j * NBIN + q] = -0.03125 * (q + 2) * (j + 1);
                k[j * NBIN + q] = 2.0 * (q + 1) + j + 0.5 * i;
            }
        }
        ffpcld(f, 1, 1, 1, n, x, &status);
        ffpcld(f, 2, 1, 1, n, y, &status);
        ffpcld(f, 3, 1, 1, n, w, &status);
        ffpcld(f, 4, 1, 1, n * NBIN, e, &status);
        ffpcld(f, 5, 1, 1, n * NBIN, b, &status);
        ffpcld(f, 6, 1, 1, n * NBIN, k, &status);
        check(status);
    }
    ffclos(f, &status);
    check(status);
    return 0;
}

static int read_file(const char *name)
{
    fitsfile *f;
    int status = 0, nhdu = 0, hdutype;
    long nfield, nbin;
    ffopen(&f, name, READONLY, &status);
    check(status);
    ffgkyj(f, "NFIELD", &nfield, NULL, &status);
    ffgkyj(f, "NBIN", &nbin, NULL, &status);
    ffthdu(f, &nhdu, &status);
    check(status);
    printf("NFIELD %ld NBIN %ld HDUS %d\n", nfield, nbin, nhdu);
    for (int i = 0; i < nfield; ++i) {
        long n;
        char *lens = NULL;
        ffmahd(f, i + 2, &hdutype, &status);
        ffgnrw(f, &n, &status);
        ffgkls(f, "LENSFILE", &lens, NULL, &status);
        check(status);
        printf("FIELD %d ROWS %ld LENSFILE %s\n", i + 1, n, lens);
        free(lens);
        const char *cols[6] = {"x", "y", "w", "xi_E_MARKS", "xi_B_MARKS", "K_MARKS"};
        for (int c = 0; c < 6; ++c) {
            int cn;
            ffgcno(f, CASEINSEN, (char *)cols[c], &cn, &status);
            const long m = c < 3 ? n : n * nbin;
            double *v = malloc(sizeof(double) * (size_t)m);
            ffgcvd(f, cn, 1, 1, m, 0.0, v, NULL, &status);
            check(status);
            printf("%s", cols[c]);
            for (long j = 0; j < m; ++j) printf(" %.17g", v[j]);
            printf("\n");
            free(v);
        }
    }
    ffclos(f, &status);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s out.fits | %s --read in.fits\n", argv[0], argv[0]);
        return 2;
    }
    if (strcmp(argv[1], "--read") == 0 && argc > 2) return read_file(argv[2]);
    return write_file(argv[1]);
}
