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
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
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
    a->block_batches = 4;
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
        else if (!strcmp(s, "--block-batches")) {
            a->block_batches = atoi(v);
            if (a->block_batches < 0 || a->block_batches > 64) {
                fprintf(stderr, "[galn] --block-batches takes 0 (whole input) to 64\n");
                return -1;
            }
        }
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

// gzip (or plain) input: a thread inflates ahead of the parser into a ring of chunks, as the
// reference's inflate thread does (src/host_io.c:79-130), so that decompression and parsing
// overlap
class InflateRing {
  public:
    static constexpr int NCH = 16;
    static constexpr int CH = 4 << 20;
    ~InflateRing() { close(); }
    bool open(const char *fn)
    {
        fp_ = gzopen(fn, "rb");
        if (!fp_) return false;
        gzbuffer(fp_, 1 << 20);
        buf_.assign(NCH, std::vector<unsigned char>(CH));
        len_.assign(NCH, 0);
        th_ = std::thread([this] { run(); });
        return true;
    }
    // the next chunk, valid until the next call (which hands it back); 0 at the end of input
    int next(const unsigned char **p)
    {
        std::unique_lock<std::mutex> lk(mu_);
        if (holding_) {
            ++tail_;
            holding_ = false;
            cv_.notify_all();
        }
        if (eof_) return 0;
        cv_.wait(lk, [&] { return head_ > tail_; });
        const int k = (int)(tail_ % NCH);
        if (len_[k] == 0) {
            eof_ = true;
            return 0;
        }
        holding_ = true;
        *p = buf_[k].data();
        return len_[k];
    }
    // stops the thread; true when everything read so far decompressed without error
    bool close()
    {
        if (!fp_) return closed_ok_;
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
        const int zc = gzclose(fp_);
        fp_ = nullptr;
        closed_ok_ = !err_ && zc == Z_OK;
        return closed_ok_;
    }
    bool decompression_error() const { return err_; }

  private:
    void run()
    {
        for (;;) {
            int k;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || head_ - tail_ < NCH; });
                if (stop_) return;
                k = (int)(head_ % NCH);
            }
            int r = gzread(fp_, buf_[k].data(), (unsigned)CH);
            std::lock_guard<std::mutex> lk(mu_);
            if (r < 0) {
                err_ = true;  // a decompression error is not taken for the end of input
                r = 0;
            }
            len_[k] = r;
            ++head_;
            cv_.notify_all();
            if (r == 0) return;
        }
    }
    gzFile fp_ = nullptr;
    std::vector<std::vector<unsigned char>> buf_;
    std::vector<int> len_;
    long head_ = 0, tail_ = 0;
    bool holding_ = false, eof_ = false, stop_ = false, err_ = false, closed_ok_ = false;
    std::mutex mu_;
    std::condition_variable cv_;
    std::thread th_;
};

struct GzStream {
    InflateRing *ring;
    const unsigned char *data = nullptr;
    int begin = 0, end = 0;
    bool eof = false;
    explicit GzStream(InflateRing *r) : ring(r) {}
    bool fill()
    {
        if (eof) return false;
        begin = 0;
        end = ring->next(&data);
        if (end == 0) eof = true;
        return end > 0;
    }
    int getc()
    {
        if (begin >= end && !fill()) return -1;
        return data[begin++];
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
                const void *nl = memchr(data + begin, '\n', (size_t)(end - begin));
                i = nl ? (int)((const unsigned char *)nl - data) : end;
            } else {
                while (i < end && !isspace(data[i])) ++i;
            }
            gotany = true;
            s.append((const char *)data + begin, (size_t)(i - begin));
            begin = i + 1;
            if (i < end) {
                if (dret) *dret = data[i];
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
    // src/host_io.c:151-205 (bwa_read_seq batches)
    InflateRing ring;
    if (!ring.open(fn)) {
        fprintf(stderr, "[galn] cannot open %s\n", fn);
        return nullptr;
    }
    GzStream ks(&ring);
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
    if (!ring.close()) {
        fprintf(stderr, "[galn] error reading %s (%s)\n", fn,
                ring.decompression_error() ? "decompression error" : "truncated or corrupt input");
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

namespace {

// A block of whole bwa batches of the input, for the streamed path: queries, lengths, batch ids
// counted from 0 within the block, names.
struct ReadBlock {
    std::vector<uint8_t> query;
    std::vector<uint64_t> qoff;
    std::vector<int32_t> qlen, batch;
    std::vector<uint64_t> name_off;
    std::vector<char> names;
    int n_batches = 0;
    int n() const { return (int)qlen.size(); }
    const char *name(int i) const { return names.data() + name_off[(size_t)i]; }
};

// Parses the input on a thread of its own into blocks of `max_batches` bwa batches, up to `ahead`
// blocks ahead of the consumer, with the record and batch rules of galn_read_fastq.
class BlockReader {
  public:
    ~BlockReader() { stop(); }
    bool open(const char *fn, int max_batches, int ahead)
    {
        if (!ring_.open(fn)) return false;
        max_batches_ = max_batches;
        ahead_ = ahead;
        th_ = std::thread([this] { run(); });
        return true;
    }
    // the next block in input order; false at the end of input (see ok())
    bool next(std::unique_ptr<ReadBlock> &b)
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !q_.empty() || done_; });
        if (q_.empty()) return false;
        b = std::move(q_.front());
        q_.pop_front();
        cv_.notify_all();
        return true;
    }
    // after the end: the input decompressed without error
    bool ok()
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return done_; });
        return ok_;
    }
    bool decompression_error() const { return ring_.decompression_error(); }
    void stop()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
    }

  private:
    void push(std::unique_ptr<ReadBlock> &b, int n_batches)
    {
        b->n_batches = n_batches;
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return (int)q_.size() < ahead_ || stop_; });
        q_.push_back(std::move(b));
        cv_.notify_all();
        b.reset(new ReadBlock);
    }
    void run()
    {
        GzStream ks(&ring_);
        Record rec;
        auto blk = std::make_unique<ReadBlock>();
        int local = 0, in_batch = 0;  // batches completed in this block; reads in the current batch
        for (;;) {
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (stop_) break;
            }
            const int l = read_record(ks, rec);
            if (l < 0) {
                // a malformed record ends the batch; an empty batch ends the input (src/host_io.c:164-186)
                if (in_batch == 0 || l == -1) break;
                in_batch = 0;
                if (++local == max_batches_) {
                    push(blk, local);
                    local = 0;
                }
                continue;
            }
            if (rec.seq.empty()) continue;
            const int len = (int)rec.seq.size();
            blk->qoff.push_back(blk->query.size());
            blk->qlen.push_back(len);
            blk->batch.push_back(local);
            const size_t o = blk->query.size();
            blk->query.resize(o + (size_t)len);
            for (int j = 0; j < len; ++j) {
                const unsigned char b = nt4_of((unsigned char)rec.seq[(size_t)(len - 1 - j)]);
                blk->query[o + (size_t)j] = b > 3 ? 4 : (uint8_t)(3 - b);
            }
            blk->name_off.push_back(blk->names.size());
            blk->names.insert(blk->names.end(), rec.name.begin(), rec.name.end());
            blk->names.push_back('\0');
            if (++in_batch == GALN_HOST_ALN_BATCH) {
                in_batch = 0;
                if (++local == max_batches_) {
                    push(blk, local);
                    local = 0;
                }
            }
        }
        if (blk->n() > 0) push(blk, local + (in_batch > 0 ? 1 : 0));
        const bool ok = ring_.close();
        std::lock_guard<std::mutex> lk(mu_);
        ok_ = ok;
        done_ = true;
        cv_.notify_all();
    }
    InflateRing ring_;
    int max_batches_ = 4, ahead_ = 3;
    std::deque<std::unique_ptr<ReadBlock>> q_;
    bool done_ = false, ok_ = false, stop_ = false;
    std::mutex mu_;
    std::condition_variable cv_;
    std::thread th_;
};

}  // namespace

// ---------------------------------------------------------------------------
// .sai

// per read in input order: int32 n_aln (an errored read: 0) and its records
static int sai_write_reads(FILE *fp, const galn_gap_opt_t *o, int32_t n, const int32_t *n_aln, const uint64_t *rec_off,
                           const galn_aln_rec_t *rec, bool header)
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
    bool ok = !header || (put("SAI\1", 4) && put(o, sizeof(*o)));
    for (int32_t i = 0; i < n && ok; ++i) {
        const int32_t na = n_aln[i] > 0 ? n_aln[i] : 0;
        ok = put(&na, 4);
        if (na > 0 && ok) ok = put(rec + rec_off[i], (size_t)na * sizeof(galn_aln_rec_t));
    }
    if (ok && !buf.empty()) ok = fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
    if (!ok || fflush(fp) != 0 || ferror(fp)) return 5;
    return 0;
}

extern "C" int galn_sai_write(FILE *fp, const galn_gap_opt_t *o, int32_t n, const int32_t *n_aln,
                              const uint64_t *rec_off, const galn_aln_rec_t *rec)
{
    return sai_write_reads(fp, o, n, n_aln, rec_off, rec, true);
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
            "         --cpu-interleave INT  reads one CPU thread advances in turn [2, 4 or 8 by index size]\n"
            "         --verify           re-run every read on the CPU and compare (exit 6 on a difference)\n"
            "         --block-batches INT  stream the input to the GPU in blocks of INT bwa batches of\n"
            "                            0x40000 reads, read, searched and written at the same time;\n"
            "                            0 reads the whole input first [4]\n"
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
        // pop traces only for --trace output and --verify (hashing every pop costs the engine time)
        galn_hip_out O = so.view(A.trace || A.verify);
        const int r = galn_cpu_search(bwt, meta, R, list ? list->data() : nullptr, n, &O, A.cpu_threads, A.cpu_interleave);
        if (r == GALN_HIP_E_RECCAP) {
            so.rec.resize(O.n_rec);
            continue;
        }
        return r;
    }
}

// The reads a GPU search left as outside its kernels' limits (GALN_READ_ERR_UNSUP), searched on the CPU
// engine and merged into `res` in read order, so the records are the same whichever engine found them.
static int finish_unsupported(const uint32_t *bwt, const galn_bwt_meta_t *meta, const galn_hip_reads *R,
                              const galn_aln_args_t &A, SearchOut &res, int *n_cpu)
{
    const int n = R->n_reads;
    std::vector<int32_t> cpu_list;
    for (int i = 0; i < n; ++i)
        if (res.n_aln[(size_t)i] == GALN_READ_ERR_UNSUP) cpu_list.push_back(i);
    *n_cpu = (int)cpu_list.size();
    if (cpu_list.empty()) return 0;
    SearchOut c;
    const int r = cpu_run(bwt, meta, R, &cpu_list, A, c);
    if (r) return r;
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
    return 0;
}

// The streamed GPU path: blocks of whole bwa batches are read and parsed on their own threads,
// searched by the serving loop (up to three launches on the device at once, so one block's
// search overlaps the next) and written in input order as they finish. Returns the exit code.
static int aln_stream(const galn_aln_args_t &A, FILE *out, const uint32_t *bwt, const galn_bwt_meta_t &meta,
                      BlockReader &br, double t0, double t_loaded)
{
    galn_hip_index *idx = nullptr;
    if (galn_hip_index_open(bwt, &meta, A.device, &idx)) {
        fprintf(stderr, "[galn] GPU: %s (use --cpu-only to run without a GPU)\n", galn_hip_last_error());
        return 10;
    }
    const bool want_trace = A.trace != 0;
    galn_hip_server_cfg_t cfg{};
    cfg.max_launch_reads = A.block_batches * GALN_HOST_ALN_BATCH;
    cfg.quantum_reads = cfg.max_launch_reads;
    cfg.max_queued_bytes = (uint64_t)64 << 30;
    cfg.flags = want_trace ? GALN_HIP_TRACE : 0u;
    cfg.form = A.form;
    cfg.slots = 3;
    galn_hip_server *S = nullptr;
    if (galn_hip_server_open(idx, &cfg, &S)) {
        fprintf(stderr, "[galn] GPU: %s\n", galn_hip_last_error());
        galn_hip_index_close(idx);
        return 10;
    }
    struct Pend {
        std::unique_ptr<ReadBlock> b;
        std::vector<galn_search_opt_t> bopts;
        std::vector<int32_t> nstacks, maxdiff;
        galn_hip_reads R{};
        uint64_t id = 0;
    };
    std::deque<Pend> pend;
    int err = 0, fatal = 0, n_cpu = 0;
    long n_reads = 0, mapped = 0, n_rec = 0, blocks = 0;
    double t_cpu = 0, pops_total = 0;
    if (sai_write_reads(out, &A.opt, 0, nullptr, nullptr, nullptr, true)) err = 5;
    // the oldest block's results, completed on the CPU where needed, onto the .sai
    auto collect_front = [&]() {
        Pend &p = pend.front();
        const int n = p.b->n();
        int r = galn_hip_server_wait(S, p.id, -1);
        SearchOut res;
        res.resize(n);
        galn_hip_out O = res.view(want_trace);
        double lat = 0;
        if (r >= 0) {
            r = galn_hip_server_result(S, p.id, &O, &lat);
            if (r == GALN_HIP_E_RECCAP) {
                res.rec.resize(O.n_rec);
                O = res.view(want_trace);
                r = galn_hip_server_result(S, p.id, &O, &lat);
            }
        }
        if (r < 0) {
            fprintf(stderr, "[galn] GPU search failed: %s\n", galn_hip_last_error());
            fatal = 10;
        } else {
            res.rec.resize(O.n_rec);
            const double tc = now_s();
            int nc = 0;
            if ((r = finish_unsupported(bwt, &meta, &p.R, A, res, &nc))) {
                fprintf(stderr, "[galn] CPU search failed (%d)\n", r);
                fatal = 10;
            }
            n_cpu += nc;
            if (nc) t_cpu += now_s() - tc;
            for (int i = 0; i < n && !fatal; ++i) {
                if (res.n_aln[(size_t)i] < 0) {
                    fprintf(stderr, "[galn] read %ld (%s): search error %d\n", n_reads + i, p.b->name(i),
                            res.n_aln[(size_t)i]);
                    err = err ? err : 4;
                } else if (res.n_aln[(size_t)i] > 0) {
                    ++mapped;
                    n_rec += res.n_aln[(size_t)i];
                }
                pops_total += (double)res.pops[(size_t)i];
            }
            // a read with an error is written with 0 records (exit 4); after a write error, nothing more
            if (!fatal && err != 5 &&
                sai_write_reads(out, &A.opt, n, res.n_aln.data(), res.rec_off.data(), res.rec.data(), false))
                err = 5;
        }
        n_reads += n;
        ++blocks;
        pend.pop_front();
    };
    std::unique_ptr<ReadBlock> b;
    while (!fatal && br.next(b)) {
        Pend p;
        p.b = std::move(b);
        const int n = p.b->n(), nb = std::max(1, p.b->n_batches);
        p.bopts.resize((size_t)nb);
        p.nstacks.resize((size_t)nb);
        p.maxdiff.resize((size_t)std::max(n, 1));
        galn_prepare_batches(&A.opt, n, p.b->qlen.data(), p.b->batch.data(), nb, p.bopts.data(), p.nstacks.data(),
                             p.maxdiff.data());
        p.R.n_reads = n;
        p.R.query = p.b->query.data();
        p.R.qoff = p.b->qoff.data();
        p.R.qlen = p.b->qlen.data();
        p.R.max_diff = p.maxdiff.data();
        p.R.batch = p.b->batch.data();
        p.R.n_batches = nb;
        p.R.bopts = p.bopts.data();
        p.R.n_stacks = p.nstacks.data();
        for (;;) {
            const int r = galn_hip_server_submit(S, &p.R, &p.id);
            if (r == GALN_HIP_E_BUSY && !pend.empty()) {
                collect_front();
                continue;
            }
            if (r) {
                fprintf(stderr, "[galn] GPU: %s\n", galn_hip_last_error());
                fatal = 10;
            }
            break;
        }
        if (fatal) break;
        pend.push_back(std::move(p));
        // keep the blocks in flight bounded; write the oldest as soon as it is done
        while (!fatal && !pend.empty() && ((int)pend.size() > 4 || galn_hip_server_wait(S, pend.front().id, 0) != 0))
            collect_front();
    }
    while (!fatal && !pend.empty()) collect_front();
    const bool input_ok = fatal ? true : br.ok();
    br.stop();
    galn_hip_server_close(S);
    galn_hip_index_close(idx);
    if (out != stdout && fclose(out) != 0 && !err) err = 5;
    if (fatal) return fatal;
    if (!input_ok) {
        fprintf(stderr, "[galn] error reading %s (%s)\n", A.reads,
                br.decompression_error() ? "decompression error" : "truncated or corrupt input");
        return 3;
    }
    const double t_end = now_s();
    fprintf(stderr, "[galn] %ld reads (%ld with hits, %ld records) in %.2f s: index load %.2f s, streamed read + search + "
                    "write %.2f s (%ld blocks)\n",
            n_reads, mapped, n_rec, t_end - t0, t_loaded - t0, t_end - t_loaded, blocks);
    if (A.verbose) {
        fprintf(stderr, "[galn] GPU %d, %s form, streamed: %.3e pops, %.3e reads/s end to end\n", A.device,
                A.form == GALN_FORM_PLAIN ? "plain" : "batched", pops_total, n_reads / std::max(t_end - t0, 1e-9));
        if (n_cpu)
            fprintf(stderr, "[galn] CPU engine: %d reads in %.3f s (%.3e reads/s)\n", n_cpu, t_cpu,
                    t_cpu > 0 ? n_cpu / t_cpu : 0.0);
    }
    return err;
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
    if (!A.cpu_only && !A.verify && A.block_batches > 0) {
        // the streamed GPU path: the input is parsed while the index loads
        BlockReader br;
        const bool opened = br.open(A.reads, A.block_batches, 3);
        galn_bwt_meta_t meta;
        uint32_t *bwt = galn_load_bwt(bwtfn.c_str(), &meta);
        if (!bwt) {
            br.stop();
            return 2;
        }
        if (!opened) {
            fprintf(stderr, "[galn] cannot open %s\n", A.reads);
            galn_free_bwt(bwt);
            return 3;
        }
        const int rc = aln_stream(A, out, bwt, meta, br, t0, now_s());
        galn_free_bwt(bwt);
        return rc;
    }
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
        const double tc = now_s();
        r = finish_unsupported(bwt, &meta, &R, A, res, &n_cpu);
        if (r) {
            fprintf(stderr, "[galn] CPU search failed (%d)\n", r);
            cleanup();
            return 10;
        }
        if (n_cpu) t_cpu = now_s() - tc;
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
