// galn_host.cpp -- host side of `galn aln` (see include/galn_host.h): option parsing,
// per-batch options, .bwt loading, FASTQ/FASTA input, .sai output, and the aln driver
// that hands the reads to the GPU search (galn_hip.h), and to the CPU engine (galn_cpu.h) for
// --cpu-only and for the reads outside the GPU kernels' limits.

#include <cctype>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include "galn_cpu.h"
#include "galn_hip.h"
#include "galn_host.h"

// ---------------------------------------------------------------------------
// options

extern "C" void galn_gap_opt_default(galn_gap_opt_t *o)
{
    // bwa 0.7.19 gap_init_opt(): src/opt_build.h:23-36
    memset(o, 0, sizeof(*o));
    o->s_mm = 3;
    o->s_gapo = 11;
    o->s_gape = 4;
    o->max_diff = -1;
    o->max_gapo = 1;
    o->max_gape = 6;
    o->indel_end_skip = 5;
    o->max_del_occ = 10;
    o->max_entries = 2000000;
    o->mode = GALN_MODE_GAPE | GALN_MODE_COMPREAD;
    o->seed_len = 32;
    o->max_seed_diff = 2;
    o->fnr = 0.04f;
    o->n_threads = 1;
    o->max_top2 = 30;
    o->trim_qual = 0;
}

// options of the reference that change nothing here (CPU/GPU sharing and tuning of its own
// engines); accepted so existing command lines keep working
static const char *k_ignored_valued[] = {"--handoff", "--handoff-when", "--gpus", "--light-cap", "--warps",
                                         "--pool-gb", "--alloc", nullptr};

extern "C" int galn_parse_aln_args(int argc, char **argv, galn_aln_args_t *a)
{
    // src/galn.cu:185-250
    int opte = -1, npos = 0;
    memset(a, 0, sizeof(*a));
    galn_gap_opt_default(&a->opt);
    galn_gap_opt_t *o = &a->opt;
    a->form = GALN_FORM_BATCHED;
    for (int i = 0; i < argc; ++i) {
        const char *s = argv[i];
        const char *v = (i + 1 < argc) ? argv[i + 1] : nullptr;
        if (s[0] != '-' || !s[1]) {
            if (npos == 0) a->prefix = s;
            else if (npos == 1) a->reads = s;
            else {
                fprintf(stderr, "[galn] unexpected argument '%s'\n", s);
                return -1;
            }
            ++npos;
            continue;
        }
        if (!strcmp(s, "-L")) { o->mode |= GALN_MODE_LOGGAP; continue; }
        if (!strcmp(s, "-N")) { o->mode |= GALN_MODE_NONSTOP; o->max_top2 = 0x7fffffff; continue; }
        if (!strcmp(s, "--gpu-only")) continue;  // the GPU always does every read here
        if (!strcmp(s, "-v")) { a->verbose = 1; continue; }
        if (!strcmp(s, "--trace")) { a->trace = 1; continue; }
        if (!strcmp(s, "--cpu-only")) { a->cpu_only = 1; continue; }
        if (!strcmp(s, "--verify")) { a->verify = 1; continue; }
        if (!v) {
            fprintf(stderr, "[galn] option %s needs a value\n", s);
            return -1;
        }
        bool ignored = false;
        for (int k = 0; k_ignored_valued[k]; ++k) ignored = ignored || !strcmp(s, k_ignored_valued[k]);
        if (ignored) { ++i; continue; }
        if (!strcmp(s, "-n")) {
            if (strstr(v, ".")) o->fnr = (float)atof(v), o->max_diff = -1;
            else o->max_diff = atoi(v), o->fnr = -1.0f;
        }
        else if (!strcmp(s, "-o")) o->max_gapo = atoi(v);
        else if (!strcmp(s, "-e")) opte = atoi(v);
        else if (!strcmp(s, "-M")) o->s_mm = atoi(v);
        else if (!strcmp(s, "-O")) o->s_gapo = atoi(v);
        else if (!strcmp(s, "-E")) o->s_gape = atoi(v);
        else if (!strcmp(s, "-d")) o->max_del_occ = atoi(v);
        else if (!strcmp(s, "-i")) o->indel_end_skip = atoi(v);
        else if (!strcmp(s, "-l")) o->seed_len = atoi(v);
        else if (!strcmp(s, "-k")) o->max_seed_diff = atoi(v);
        else if (!strcmp(s, "-m")) o->max_entries = atoi(v);
        else if (!strcmp(s, "-dt") || !strcmp(s, "-t")) o->n_threads = atoi(v);  // header value only
        else if (!strcmp(s, "-R")) o->max_top2 = atoi(v);
        else if (!strcmp(s, "-f")) a->out = v;
        else if (!strcmp(s, "-q")) {
            if (atoi(v) >= 1) {
                fprintf(stderr, "[galn] -q (quality trimming) is not supported\n");
                return -1;
            }
        }
        else if (!strcmp(s, "--device")) a->device = atoi(v);
        else if (!strcmp(s, "--cpu-workers")) a->cpu_threads = atoi(v);
        else if (!strcmp(s, "--cpu-interleave")) a->cpu_interleave = atoi(v);
        else if (!strcmp(s, "--form")) {
            if (!strcmp(v, "plain")) a->form = GALN_FORM_PLAIN;
            else if (!strcmp(v, "batched")) a->form = GALN_FORM_BATCHED;
            else {
                fprintf(stderr, "[galn] --form must be plain or batched\n");
                return -1;
            }
        }
        else {
            fprintf(stderr, "[galn] unsupported option %s\n", s);
            return -1;
        }
        ++i;
    }
    if (opte > 0) {
        o->max_gape = opte;
        o->mode &= ~GALN_MODE_GAPE;
    }
    if (o->n_threads < 1) o->n_threads = 1;
    if (npos != 2) return -1;
    return 0;
}

extern "C" int galn_cal_maxdiff(int l, double err, double thres)
{
    // bwa_cal_maxdiff, src/aln_host.c:9-21 (this file is compiled with -ffp-contract=off)
    double elambda = exp(-l * err);
    double sum, y = 1.0;
    int k, x = 1;
    for (k = 1, sum = elambda; k < 1000; ++k) {
        y *= l * err;
        x = (int)((unsigned)x * (unsigned)k);
        sum += elambda * y / x;
        if (1.0 - sum < thres) return k;
    }
    return 2;
}

extern "C" int galn_batch_opt(const galn_gap_opt_t *o, int max_len, galn_search_opt_t *b)
{
    // src/aln_host.c:23-38
    b->s_mm = o->s_mm;
    b->s_gapo = o->s_gapo;
    b->s_gape = o->s_gape;
    b->mode = o->mode;
    b->indel_end_skip = o->indel_end_skip;
    b->max_del_occ = o->max_del_occ;
    b->max_entries = o->max_entries;
    b->max_diff = o->max_diff;
    b->max_gapo = o->max_gapo;
    b->max_gape = o->max_gape;
    b->max_seed_diff = o->max_seed_diff;
    b->seed_len = o->seed_len;
    b->max_top2 = o->max_top2;
    if (o->fnr > 0.0) b->max_diff = galn_cal_maxdiff(max_len, 0.02, o->fnr);
    if (b->max_diff < b->max_gapo) b->max_gapo = b->max_diff;
    return (b->max_diff + 1) * b->s_mm + (b->max_gapo + 1) * b->s_gapo + (b->max_gape + 1) * b->s_gape;
}

extern "C" int galn_read_maxdiff(const galn_gap_opt_t *o, const galn_search_opt_t *b, int len)
{
    return (o->fnr > 0.0) ? galn_cal_maxdiff(len, 0.02, o->fnr) : b->max_diff;
}

extern "C" int galn_prepare_batches(const galn_gap_opt_t *o, int n, const int32_t *qlen, const int32_t *batch,
                                    int max_batches, galn_search_opt_t *bopts, int32_t *n_stacks, int32_t *max_diff)
{
    // src/galn.cu:605-621: options per bwa batch from its longest read, max_diff per read
    int nb = 0;
    int i = 0;
    if (n == 0) {
        if (max_batches < 1) return -1;
        n_stacks[0] = galn_batch_opt(o, 0, &bopts[0]);
        return 1;
    }
    while (i < n) {
        int j = i, ml = 0;
        while (j < n && batch[j] == batch[i]) {
            if (qlen[j] > ml) ml = qlen[j];
            ++j;
        }
        if (batch[i] != nb || nb >= max_batches) return -1;  // batch ids must be 0,1,2,... in order
        n_stacks[nb] = galn_batch_opt(o, ml, &bopts[nb]);
        // cal_maxdiff depends on the length only: memoise per batch
        std::vector<int> memo;
        for (int t = i; t < j; ++t) {
            if (o->fnr > 0.0) {
                const int L = qlen[t];
                if ((int)memo.size() <= L) memo.resize(L + 1, INT_MIN);
                if (memo[L] == INT_MIN) memo[L] = galn_cal_maxdiff(L, 0.02, o->fnr);
                max_diff[t] = memo[L];
            } else {
                max_diff[t] = bopts[nb].max_diff;
            }
        }
        ++nb;
        i = j;
    }
    return nb;
}

// ---------------------------------------------------------------------------
// .bwt

extern "C" uint32_t *galn_load_bwt(const char *fn, galn_bwt_meta_t *m)
{
    // src/host_io.c:29-76: 40-byte header (primary, L2[1..4]) then the u32 body
    int fd = open(fn, O_RDONLY);
    if (fd < 0) {
        fprintf(stderr, "[galn] cannot open %s\n", fn);
        return nullptr;
    }
    struct stat st;
    uint64_t hdr[5];
    if (fstat(fd, &st) != 0 || (uint64_t)st.st_size < sizeof(hdr) || pread(fd, hdr, sizeof(hdr), 0) != (ssize_t)sizeof(hdr)) {
        fprintf(stderr, "[galn] short read on %s\n", fn);
        close(fd);
        return nullptr;
    }
    m->primary = hdr[0];
    m->L2[0] = 0;
    for (int i = 0; i < 4; ++i) m->L2[i + 1] = hdr[i + 1];
    m->seq_len = m->L2[4];
    m->bwt_size = ((uint64_t)st.st_size - sizeof(hdr)) >> 2;
    const size_t bytes = (size_t)m->bwt_size * 4, align = (size_t)2 << 20;
    const size_t alloc = std::max(align, (bytes + align - 1) / align * align);
    void *mem = nullptr;
    if (posix_memalign(&mem, align, alloc) != 0) {
        close(fd);
        return nullptr;
    }
    madvise(mem, alloc, MADV_HUGEPAGE);
    // parallel reads of 64 MB pieces
    const size_t piece = (size_t)64 << 20;
    const size_t npiece = (bytes + piece - 1) / piece;
    const unsigned nthr = std::max(1u, std::min<unsigned>(16, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    std::vector<int> ok(nthr, 1);
    for (unsigned t = 0; t < nthr; ++t)
        th.emplace_back([&, t] {
            for (size_t c = t; c < npiece; c += nthr) {
                const size_t off = c * piece, len = std::min(piece, bytes - off);
                size_t done = 0;
                while (done < len) {
                    ssize_t r = pread(fd, (char *)mem + off + done, len - done, (off_t)(sizeof(hdr) + off + done));
                    if (r <= 0) {
                        ok[t] = 0;
                        return;
                    }
                    done += (size_t)r;
                }
            }
        });
    for (auto &t : th) t.join();
    close(fd);
    for (int v : ok)
        if (!v) {
            fprintf(stderr, "[galn] short read on %s\n", fn);
            free(mem);
            return nullptr;
        }
    return (uint32_t *)mem;
}

extern "C" void galn_free_bwt(uint32_t *bwt) { free(bwt); }

// ---------------------------------------------------------------------------
// FASTQ / FASTA input with the record semantics of bwa's kseq_read (src/kseq.h:175-215)

namespace {

struct GzStream {
    gzFile fp = nullptr;
    std::vector<unsigned char> buf;
    int begin = 0, end = 0;
    bool eof = false;
    bool error = false;
    explicit GzStream(gzFile f) : fp(f), buf(1 << 20) {}
    bool fill()
    {
        if (eof) return false;
        const int r = gzread(fp, buf.data(), (unsigned)buf.size());
        if (r < 0) error = true;
        begin = 0;
        end = r > 0 ? r : 0;
        if (end == 0) eof = true;
        return end > 0;
    }
    int getc()
    {
        if (begin >= end && !fill()) return -1;
        return buf[begin++];
    }
    // append bytes up to (not including) the delimiter; the delimiter is consumed.
    // mode 0: whitespace, 1: newline. Returns -1 if nothing was read at the end of input,
    // and sets *dret to the delimiter (0 at the end of input).
    long getuntil(int mode, std::string &s, int *dret, bool append)
    {
        bool gotany = false;
        if (dret) *dret = 0;
        if (!append) s.clear();
        for (;;) {
            if (begin >= end && !fill()) break;
            int i = begin;
            if (mode == 1) {
                while (i < end && buf[i] != '\n') ++i;
            } else {
                while (i < end && !isspace(buf[i])) ++i;
            }
            gotany = true;
            s.append((const char *)buf.data() + begin, (size_t)(i - begin));
            begin = i + 1;
            if (i < end) {
                if (dret) *dret = buf[i];
                break;
            }
        }
        if (!gotany && eof && begin >= end) return -1;
        if (mode == 1 && s.size() > 1 && s.back() == '\r') s.pop_back();
        return (long)s.size();
    }
};

struct Record {
    std::string name, comment, seq, qual;
    int last_char = 0;
};

// >= 0: sequence length; -1: end of input; -2: truncated or mismatched quality
int read_record(GzStream &ks, Record &r)
{
    int c;
    if (r.last_char == 0) {
        while ((c = ks.getc()) != -1 && c != '>' && c != '@') {
        }
        if (c == -1) return -1;
        r.last_char = c;
    }
    r.comment.clear();
    r.seq.clear();
    r.qual.clear();
    if (ks.getuntil(0, r.name, &c, false) < 0) return -1;
    if (c != '\n') ks.getuntil(1, r.comment, nullptr, false);
    while ((c = ks.getc()) != -1 && c != '>' && c != '+' && c != '@') {
        if (c == '\n') continue;
        r.seq.push_back((char)c);
        ks.getuntil(1, r.seq, nullptr, true);
    }
    if (c == '>' || c == '@') r.last_char = c;
    if (c != '+') return (int)r.seq.size();
    while ((c = ks.getc()) != -1 && c != '\n') {
    }
    if (c == -1) return -2;
    while (ks.getuntil(1, r.qual, nullptr, true) >= 0 && r.qual.size() < r.seq.size()) {
    }
    r.last_char = 0;
    if (r.seq.size() != r.qual.size()) return -2;
    return (int)r.seq.size();
}

// nt4: A/C/G/T either case -> 0..3, anything else 4 (src/host_io.c:16-25)
unsigned char nt4_of(unsigned char ch)
{
    switch (ch) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

}  // namespace

struct galn_readbuf {
    std::vector<uint8_t> query;
    std::vector<uint64_t> qoff;
    std::vector<int32_t> qlen, batch;
    std::vector<uint64_t> name_off;
    std::vector<char> names;
    int32_t n_batches = 0;
};

extern "C" galn_readbuf_t *galn_read_fastq(const char *fn)
{
    // src/host_io.c:188-224 (bwa_read_seq batches)
    gzFile fp = gzopen(fn, "rb");
    if (!fp) {
        fprintf(stderr, "[galn] cannot open %s\n", fn);
        return nullptr;
    }
    gzbuffer(fp, 1 << 20);
    GzStream ks(fp);
    Record rec;
    galn_readbuf *rb = new galn_readbuf;
    int batch = 0, in_batch = 0;
    for (;;) {
        const int l = read_record(ks, rec);
        if (l < 0) {
            if (in_batch == 0 || l == -1) break;
            ++batch;
            in_batch = 0;
            continue;
        }
        if (rec.seq.empty()) continue;
        const int len = (int)rec.seq.size();
        rb->qoff.push_back(rb->query.size());
        rb->qlen.push_back(len);
        rb->batch.push_back(batch);
        const size_t o = rb->query.size();
        rb->query.resize(o + (size_t)len);
        for (int j = 0; j < len; ++j) {
            const unsigned char b = nt4_of((unsigned char)rec.seq[(size_t)(len - 1 - j)]);
            rb->query[o + (size_t)j] = b > 3 ? 4 : (uint8_t)(3 - b);
        }
        rb->name_off.push_back(rb->names.size());
        rb->names.insert(rb->names.end(), rec.name.begin(), rec.name.end());
        rb->names.push_back('\0');
        if (++in_batch == GALN_HOST_ALN_BATCH) {
            ++batch;
            in_batch = 0;
        }
    }
    const int zc = gzclose(fp);
    if (ks.error || zc != Z_OK) {
        fprintf(stderr, "[galn] error reading %s (%s)\n", fn, ks.error ? "decompression error" : "truncated or corrupt input");
        delete rb;
        return nullptr;
    }
    rb->n_batches = rb->batch.empty() ? 0 : rb->batch.back() + 1;
    return rb;
}

extern "C" int32_t galn_readbuf_n(const galn_readbuf_t *r) { return (int32_t)r->qlen.size(); }
extern "C" int32_t galn_readbuf_n_batches(const galn_readbuf_t *r) { return r->n_batches; }
extern "C" const uint8_t *galn_readbuf_query(const galn_readbuf_t *r) { return r->query.data(); }
extern "C" const uint64_t *galn_readbuf_qoff(const galn_readbuf_t *r) { return r->qoff.data(); }
extern "C" const int32_t *galn_readbuf_qlen(const galn_readbuf_t *r) { return r->qlen.data(); }
extern "C" const int32_t *galn_readbuf_batch(const galn_readbuf_t *r) { return r->batch.data(); }
extern "C" const char *galn_readbuf_name(const galn_readbuf_t *r, int32_t i)
{
    return (i >= 0 && (size_t)i < r->name_off.size()) ? r->names.data() + r->name_off[(size_t)i] : nullptr;
}
extern "C" void galn_readbuf_free(galn_readbuf_t *r) { delete r; }

// ---------------------------------------------------------------------------
// .sai

extern "C" int galn_sai_write(FILE *fp, const galn_gap_opt_t *o, int32_t n, const int32_t *n_aln,
                              const uint64_t *rec_off, const galn_aln_rec_t *rec)
{
    // src/sai_writer.h:9-21, src/galn.cu:1250-1263
    static_assert(sizeof(galn_gap_opt_t) == 64, "gap_opt_t is 64 bytes");
    static_assert(sizeof(galn_aln_rec_t) == 24, "bwt_aln1_t is 24 bytes");
    std::vector<char> buf;
    buf.reserve(1 << 22);
    auto put = [&](const void *p, size_t b) {
        buf.insert(buf.end(), (const char *)p, (const char *)p + b);
        if (buf.size() >= (1 << 22)) {
            if (fwrite(buf.data(), 1, buf.size(), fp) != buf.size()) return false;
            buf.clear();
        }
        return true;
    };
    bool ok = put("SAI\1", 4) && put(o, sizeof(*o));
    for (int32_t i = 0; i < n && ok; ++i) {
        const int32_t na = n_aln[i] > 0 ? n_aln[i] : 0;
        ok = put(&na, 4);
        if (na > 0 && ok) ok = put(rec + rec_off[i], (size_t)na * sizeof(galn_aln_rec_t));
    }
    if (ok && !buf.empty()) ok = fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
    if (!ok || fflush(fp) != 0 || ferror(fp)) return 5;
    return 0;
}

// ---------------------------------------------------------------------------
// the aln driver

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int usage()
{
    fprintf(stderr,
            "\nProgram: galn (BWA-ALN backtracking search on AMD Instinct MI355X), version " GALN_VERSION "\n"
            "         byte-identical .sai to `bwa aln` (BWA 0.7.19)\n\n"
            "Usage:   galn aln [options] <prefix> <in.fq[.gz]>\n\n"
            "bwa aln options (same meaning and defaults as BWA 0.7.19):\n"
            "         -n NUM    max #diff (int) or missing prob under 0.02 err rate (float) [0.04]\n"
            "         -o INT    maximum number of gap opens [1]\n"
            "         -e INT    maximum number of gap extensions, -1 for disabling long gaps [-1]\n"
            "         -i INT    do not put an indel within INT bp towards the ends [5]\n"
            "         -d INT    maximum occurrences for extending a long deletion [10]\n"
            "         -l INT    seed length [32]\n"
            "         -k INT    maximum differences in the seed [2]\n"
            "         -m INT    maximum entries in the queue [2000000]\n"
            "         -dt INT   thread count written to the .sai header, as bwa aln -t [1]\n"
            "         -M INT    mismatch penalty [3]\n"
            "         -O INT    gap open penalty [11]\n"
            "         -E INT    gap extension penalty [4]\n"
            "         -R INT    stop searching when there are >INT equally best hits [30]\n"
            "         -L        log-scaled gap penalty for long deletions\n"
            "         -N        non-iterative mode: search for all n-difference hits (slooow)\n"
            "         -f FILE   file to write output to instead of stdout\n"
            "         (not supported: -q, -b/-0/-1/-2, -I, -Y, -B)\n\n"
            "GPU / CPU options:\n"
            "         --device INT       HIP device to use [0]\n"
            "         --form STR         plain (one launch per 0x40000 reads) or batched (persistent\n"
            "                            kernel with a device work queue) [batched]\n"
            "         --cpu-only         search every read on the CPU (no GPU)\n"
            "         --gpu-only         the GPU searches every read it can; the CPU only those outside\n"
            "                            the GPU kernels' limits (the default here)\n"
            "         --cpu-workers INT  CPU search threads [all hardware threads]\n"
            "         --cpu-interleave INT  reads one CPU thread advances in turn [8]\n"
            "         --verify           re-run every read on the CPU and compare (exit 6 on a difference)\n"
            "         -v                 print timing to stderr\n\n");
    return 1;
}

// Records of a search: per read n_aln, offset, pops, trace; the record array.
struct SearchOut {
    std::vector<int32_t> n_aln;
    std::vector<uint64_t> rec_off, pops, trace;
    std::vector<galn_aln_rec_t> rec;
    void resize(int n)
    {
        const size_t m = (size_t)std::max(n, 1);
        n_aln.assign(m, 0);
        rec_off.assign(m, 0);
        pops.assign(m, 0);
        trace.assign(m, 0);
        if (rec.size() < 2 * m) rec.resize(2 * m);
    }
    galn_hip_out view(bool want_trace)
    {
        galn_hip_out O;
        O.n_aln = n_aln.data();
        O.rec_off = rec_off.data();
        O.rec = rec.data();
        O.rec_cap = rec.size();
        O.n_rec = 0;
        O.n_pops = pops.data();
        O.trace = want_trace ? trace.data() : nullptr;
        O.occ_bytes = nullptr;
        return O;
    }
};

// The CPU engine over `list` (or every read), growing the record array when it is too small.
static int cpu_run(const uint32_t *bwt, const galn_bwt_meta_t *meta, const galn_hip_reads *R, const std::vector<int32_t> *list,
                   const galn_aln_args_t &A, SearchOut &so)
{
    const int n = list ? (int)list->size() : R->n_reads;
    so.resize(n);
    for (;;) {
        galn_hip_out O = so.view(true);
        const int r = galn_cpu_search(bwt, meta, R, list ? list->data() : nullptr, n, &O, A.cpu_threads, A.cpu_interleave);
        if (r == GALN_HIP_E_RECCAP) {
            so.rec.resize(O.n_rec);
            continue;
        }
        return r;
    }
}

extern "C" int galn_aln_main(int argc, char **argv)
{
    galn_aln_args_t A;
    if (galn_parse_aln_args(argc, argv, &A) < 0) return usage();
    const double t0 = now_s();
    FILE *out = stdout;
    if (A.out && !(out = fopen(A.out, "wb"))) {
        fprintf(stderr, "[galn] cannot write %s\n", A.out);
        return 1;
    }
    std::string bwtfn = std::string(A.prefix) + ".bwt";
    galn_readbuf_t *reads = nullptr;
    std::thread fq([&] { reads = galn_read_fastq(A.reads); });
    galn_bwt_meta_t meta;
    uint32_t *bwt = galn_load_bwt(bwtfn.c_str(), &meta);
    fq.join();
    if (!bwt) {
        if (reads) galn_readbuf_free(reads);
        return 2;
    }
    if (!reads) {
        galn_free_bwt(bwt);
        return 3;
    }
    const double t_loaded = now_s();
    const int n = galn_readbuf_n(reads);
    const int nb = std::max(1, galn_readbuf_n_batches(reads));
    std::vector<galn_search_opt_t> bopts(nb);
    std::vector<int32_t> nstacks(nb), maxdiff(std::max(n, 1));
    galn_prepare_batches(&A.opt, n, galn_readbuf_qlen(reads), galn_readbuf_batch(reads), nb, bopts.data(), nstacks.data(),
                         maxdiff.data());
    galn_hip_reads R;
    R.n_reads = n;
    R.query = galn_readbuf_query(reads);
    R.qoff = galn_readbuf_qoff(reads);
    R.qlen = galn_readbuf_qlen(reads);
    R.max_diff = maxdiff.data();
    R.batch = galn_readbuf_batch(reads);
    R.n_batches = nb;
    R.bopts = bopts.data();
    R.n_stacks = nstacks.data();
    auto cleanup = [&]() {
        galn_free_bwt(bwt);
        galn_readbuf_free(reads);
    };

    SearchOut res;
    double kfirst = 0, ktotal = 0, t_cpu = 0;
    int n_cpu = 0;
    int r = 0;
    if (A.cpu_only) {
        // every read on the CPU engine (src/galn.cu:450-455: no device is opened)
        const double tc = now_s();
        r = cpu_run(bwt, &meta, &R, nullptr, A, res);
        t_cpu = now_s() - tc;
        n_cpu = n;
        if (r) {
            fprintf(stderr, "[galn] CPU search failed (%d)\n", r);
            cleanup();
            return 10;
        }
    } else {
        galn_hip_index *idx = nullptr;
        r = galn_hip_index_open(bwt, &meta, A.device, &idx);
        if (r) {
            fprintf(stderr, "[galn] GPU: %s (use --cpu-only to run without a GPU)\n", galn_hip_last_error());
            cleanup();
            return 10;
        }
        res.resize(n);
        const bool want_trace = A.trace || A.verify;
        galn_hip_out O = res.view(want_trace);
        const uint32_t flags = want_trace ? GALN_HIP_TRACE : 0u;
        galn_hip_readset *rs = nullptr;
        r = galn_hip_readset_upload(idx, &R, &rs);
        if (!r) r = galn_hip_readset_search(idx, rs, A.form, flags, nullptr);
        if (!r) {
            r = galn_hip_readset_results(rs, &O);
            if (r == GALN_HIP_E_RECCAP) {
                res.rec.resize(O.n_rec);
                O = res.view(want_trace);
                r = galn_hip_readset_results(rs, &O);
            }
        }
        if (rs) galn_hip_readset_kernel_ms(rs, &kfirst, &ktotal);
        if (r) {
            fprintf(stderr, "[galn] GPU search failed: %s\n", galn_hip_last_error());
            galn_hip_readset_free(rs);
            galn_hip_index_close(idx);
            cleanup();
            return 10;
        }
        galn_hip_readset_free(rs);
        galn_hip_index_close(idx);
        res.rec.resize(O.n_rec);
        // the reads outside the kernels' limits finish on the CPU engine, as the reference hands
        // them to its CPU workers (src/gapw.cuh:753-755, src/galn.cu:1151-1180)
        std::vector<int32_t> cpu_list;
        for (int i = 0; i < n; ++i)
            if (res.n_aln[(size_t)i] == GALN_READ_ERR_UNSUP) cpu_list.push_back(i);
        if (!cpu_list.empty()) {
            const double tc = now_s();
            SearchOut c;
            r = cpu_run(bwt, &meta, &R, &cpu_list, A, c);
            if (r) {
                fprintf(stderr, "[galn] CPU search failed (%d)\n", r);
                cleanup();
                return 10;
            }
            // merge: records of every read in read order, from whichever engine searched it
            std::vector<galn_aln_rec_t> rec;
            rec.reserve(res.rec.size() + c.rec.size());
            size_t j = 0;
            for (int i = 0; i < n; ++i) {
                const bool on_cpu = j < cpu_list.size() && cpu_list[j] == i;
                const galn_aln_rec_t *src = on_cpu ? c.rec.data() + c.rec_off[j] : res.rec.data() + res.rec_off[(size_t)i];
                if (on_cpu) {
                    res.n_aln[(size_t)i] = c.n_aln[j];
                    res.pops[(size_t)i] = c.pops[j];
                    res.trace[(size_t)i] = c.trace[j];
                    ++j;
                }
                res.rec_off[(size_t)i] = rec.size();
                if (res.n_aln[(size_t)i] > 0) rec.insert(rec.end(), src, src + res.n_aln[(size_t)i]);
            }
            res.rec.swap(rec);
            n_cpu = (int)cpu_list.size();
            t_cpu = now_s() - tc;
        }
    }
    const double t_search = now_s();

    int verify_fail = 0;
    if (A.verify && !A.cpu_only) {
        // every GPU read again on the CPU engine: records, pops and traces must agree
        // (src/galn.cu:1222-1248 re-runs the reads of its warp engine the same way)
        SearchOut c;
        r = cpu_run(bwt, &meta, &R, nullptr, A, c);
        long bad = 0;
        for (int i = 0; i < n && !r; ++i) {
            const int na = res.n_aln[(size_t)i];
            bool ok = c.n_aln[(size_t)i] == na && c.pops[(size_t)i] == res.pops[(size_t)i] &&
                      c.trace[(size_t)i] == res.trace[(size_t)i];
            if (ok && na > 0)
                ok = !memcmp(c.rec.data() + c.rec_off[(size_t)i], res.rec.data() + res.rec_off[(size_t)i],
                             (size_t)na * sizeof(galn_aln_rec_t));
            if (!ok && bad++ < 10)
                fprintf(stderr, "[galn verify] read %d (%s) differs: cpu n=%d pops=%llu | gpu n=%d pops=%llu\n", i,
                        galn_readbuf_name(reads, i), c.n_aln[(size_t)i], (unsigned long long)c.pops[(size_t)i], na,
                        (unsigned long long)res.pops[(size_t)i]);
        }
        if (r) bad = 1;
        fprintf(stderr, "[galn verify] %d reads re-run on the CPU: %ld differ -> %s\n", n, bad, bad ? "FAIL" : "PASS");
        verify_fail = bad ? 6 : 0;
    }

    int err = 0;
    long mapped = 0, n_rec = 0;
    for (int i = 0; i < n; ++i) {
        if (res.n_aln[(size_t)i] < 0) {
            fprintf(stderr, "[galn] read %d (%s): search error %d\n", i, galn_readbuf_name(reads, i), res.n_aln[(size_t)i]);
            err = 4;
        } else if (res.n_aln[(size_t)i] > 0) {
            ++mapped;
            n_rec += res.n_aln[(size_t)i];
        }
    }
    if (galn_sai_write(out, &A.opt, n, res.n_aln.data(), res.rec_off.data(), res.rec.data())) {
        fprintf(stderr, "[galn] error writing the .sai\n");
        err = 5;
    }
    if (out != stdout && fclose(out) != 0) err = 5;
    const double t_end = now_s();
    fprintf(stderr, "[galn] %d reads (%ld with hits, %ld records) in %.2f s: load %.2f s, search %.2f s, write %.2f s\n", n,
            mapped, n_rec, t_end - t0, t_loaded - t0, t_search - t_loaded, t_end - t_search);
    if (A.verbose) {
        double pops_total = 0;
        for (int i = 0; i < n; ++i) pops_total += (double)res.pops[(size_t)i];
        if (!A.cpu_only)
            fprintf(stderr, "[galn] GPU %d, %s form: kernels %.3f s (first tier %.3f s), %.3e pops, %.3e reads/s\n",
                    A.device, A.form == GALN_FORM_PLAIN ? "plain" : "batched", ktotal / 1e3, kfirst / 1e3, pops_total,
                    ktotal > 0 ? n / (ktotal / 1e3) : 0.0);
        if (n_cpu)
            fprintf(stderr, "[galn] CPU engine: %d reads in %.3f s (%.3e reads/s)\n", n_cpu, t_cpu,
                    t_cpu > 0 ? n_cpu / t_cpu : 0.0);
        if (A.trace)
            for (int i = 0; i < n && i < 5; ++i)
                fprintf(stderr, "[galn] read %d: n_aln %d pops %llu trace %016llx\n", i, res.n_aln[(size_t)i],
                        (unsigned long long)res.pops[(size_t)i], (unsigned long long)res.trace[(size_t)i]);
    }
    cleanup();
    return err ? err : verify_fail;
}
