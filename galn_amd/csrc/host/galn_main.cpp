// galn_main.cpp -- the `galn` executable: `galn aln ...` and `galn --version`
// (reference src/galn.cu:1294-1304; index/samse/sampe are not part of this build).
#include <cstdio>
#include <cstring>

#include "galn_host.h"

int main(int argc, char **argv)
{
    if (argc >= 2 && !strcmp(argv[1], "aln")) return galn_aln_main(argc - 2, argv + 2);
    if (argc >= 2 && (!strcmp(argv[1], "--version") || !strcmp(argv[1], "version"))) {
        printf("G-ALN %s\n", GALN_VERSION);
        return 0;
    }
    if (argc >= 2) fprintf(stderr, "[galn] unknown command '%s'\n", argv[1]);
    fprintf(stderr, "Usage: galn aln [options] <prefix> <in.fq[.gz]>   (galn aln for the option list)\n");
    return 1;
}
