// CPU side of the Mars encode path above the search provider (see
// include/mars_codec.h): entropy models and the stack rANS coder, the
// `.mars` v0 writer, the frozen rate snapshot, the threshold and RD quadtree
// walks with the four leaf modes.
//
// Every binary64 expression keeps the reference's operation order and is
// compiled with -ffp-contract=off, so values round exactly as the reference's
// Rust does (Rust never fuses a*b+c).  Transcendentals (log2 for rates, cos
// for the mode-3 DCT) come from the C library, as Rust's f64 methods do.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <unordered_map>
#include <vector>

#include "mars_codec.h"

namespace {

// ---------------------------------------------------------------- geometry (ifs.rs:12-117)

uint32_t ceil_log2(uint32_t q) { return q <= 1 ? 0 : 32 - (uint32_t)__builtin_clz(q - 1); }

struct Header {
    uint32_t bits_alfa, bits_beta, min_size, max_size, shift, width, height, int_max_alfa;
    uint32_t qstep_bits; // ResidualQstep: binary32 bits (quant.rs:18-64)
    uint32_t virtual_size() const { return 1u << ceil_log2(std::max(width, height)); }
    uint32_t bits_coord_row() const { return ceil_log2(height / shift); }
    uint32_t bits_coord_col() const { return ceil_log2(width / shift); }
    double qstep() const
    {
        float f;
        memcpy(&f, &qstep_bits, 4);
        return (double)f;
    }
};

struct Leaf {
    uint32_t row, col, size;
    uint8_t mode;
    uint32_t qalfa, qbeta;
    uint8_t isometry;
    uint32_t dom_row, dom_col;
    int32_t qgx, qgy;
    std::vector<int32_t> residual;
};

Leaf dc_leaf(uint32_t row, uint32_t col, uint32_t size, uint32_t qbeta)
{
    Leaf l{};
    l.row = row;
    l.col = col;
    l.size = size;
    l.mode = 0;
    l.qbeta = qbeta;
    return l;
}

// isometry.rs:7-19
void iso_map(uint32_t k, uint32_t u, uint32_t v, uint32_t size, uint32_t *i, uint32_t *j)
{
    switch (k) {
    case 0: *i = u; *j = v; break;
    case 1: *i = size - 1 - v; *j = u; break;
    case 2: *i = v; *j = size - 1 - u; break;
    case 3: *i = size - 1 - u; *j = size - 1 - v; break;
    case 4: *i = u; *j = size - 1 - v; break;
    case 5: *i = size - 1 - u; *j = v; break;
    case 6: *i = v; *j = u; break;
    default: *i = size - 1 - v; *j = size - 1 - u; break;
    }
}

// ---------------------------------------------------------------- entropy (mars-entropy/src/lib.rs)

constexpr int PRECISION = 24;
constexpr uint64_t PRECISION_TOTAL = 1ull << PRECISION;
constexpr uint64_t RESCALE_THRESHOLD = 1ull << 16;

// lib.rs:36-121: order-0 adaptive model, Laplace-seeded, +32 per symbol,
// halved (floor 1) when the total reaches 2^16; fixed-point CDF by integer
// division with one reserved slot per symbol.
struct AdaptiveModel {
    std::vector<uint32_t> counts;
    uint64_t total;
    explicit AdaptiveModel(uint32_t alphabet) : counts(alphabet, 1), total(alphabet) {}
    void update(uint32_t s)
    {
        counts[s] += 32;
        total += 32;
        if (total >= RESCALE_THRESHOLD) {
            uint64_t t = 0;
            for (auto &c : counts) {
                c = std::max<uint32_t>(c >> 1, 1);
                t += c;
            }
            total = t;
        }
    }
    uint32_t scaled(uint64_t c) const
    {
        const uint64_t free_weight = PRECISION_TOTAL - counts.size();
        return (uint32_t)(((unsigned __int128)c * free_weight) / total);
    }
    void left_and_prob(uint32_t symbol, uint32_t *left, uint32_t *prob) const
    {
        uint64_t before = 0;
        for (uint32_t i = 0; i < symbol; i++) before += counts[i];
        const uint64_t after = before + counts[symbol];
        const uint32_t l = scaled(before) + symbol;
        const uint32_t r = scaled(after) + symbol + 1;
        *left = l;
        *prob = r - l;
    }
    double bits_for(uint32_t symbol) const
    {
        uint32_t l, p;
        left_and_prob(symbol, &l, &p);
        return -std::log2((double)p / (double)PRECISION_TOTAL);
    }
    // lib.rs:104-121: the symbol whose fixed-point interval holds `quantile`
    // (< 2^24, which the last symbol's right end always reaches).
    void symbol_for_quantile(uint32_t quantile, uint32_t *sym, uint32_t *left, uint32_t *prob) const
    {
        uint64_t cum = 0;
        const uint32_t n = (uint32_t)counts.size();
        for (uint32_t i = 0; i < n; i++) {
            const uint64_t next = cum + counts[i];
            const uint32_t l = scaled(cum) + i, r = scaled(next) + i + 1;
            if (quantile < r || i + 1 == n) {
                *sym = i;
                *left = l;
                *prob = r - l;
                return;
            }
            cum = next;
        }
    }
};

typedef std::pair<uint8_t, uint32_t> ContextKey;

struct Event {
    ContextKey ctx;
    uint32_t alphabet, symbol;
};

typedef std::map<ContextKey, AdaptiveModel> Models;

// Stack rANS with 64-bit state and 32-bit words, PRECISION 24 -- the
// published algorithm of constriction 0.5's AnsCoder<u32, u64> (the
// reference's DefaultAnsCoder, lib.rs:280-290): an empty coder has state 0;
// before encoding a symbol of probability p, if state >> 40 >= p the low word
// is pushed to the bulk and the state shifted right by 32; then
// state = (state / p) << 24 | (left + state % p).  Symbols are pushed in
// reverse of decode order.  The compressed words are the bulk followed by the
// state's nonzero words, least significant first (a decoder pops from the
// end, so the top word comes out first); bytes are little-endian words
// (lib.rs:337-343).
std::vector<uint8_t> ans_encode(const std::vector<std::pair<uint32_t, uint32_t>> &rec)
{
    uint64_t state = 0;
    std::vector<uint32_t> words;
    for (auto it = rec.rbegin(); it != rec.rend(); ++it) {
        const uint64_t left = it->first, prob = it->second;
        if ((state >> (64 - PRECISION)) >= prob) {
            words.push_back((uint32_t)state);
            state >>= 32;
        }
        const uint64_t rem = state % prob, prefix = state / prob;
        state = (prefix << PRECISION) | (left + rem);
    }
    if (state != 0) {
        words.push_back((uint32_t)state);
        if (state >> 32) words.push_back((uint32_t)(state >> 32));
    }
    std::vector<uint8_t> out(words.size() * 4);
    for (size_t i = 0; i < words.size(); i++) {
        out[4 * i] = (uint8_t)words[i];
        out[4 * i + 1] = (uint8_t)(words[i] >> 8);
        out[4 * i + 2] = (uint8_t)(words[i] >> 16);
        out[4 * i + 3] = (uint8_t)(words[i] >> 24);
    }
    return out;
}

// lib.rs:240-260 record + 280-290 encode_recorded.
std::vector<uint8_t> entropy_encode(const std::vector<Event> &events)
{
    Models models;
    std::vector<std::pair<uint32_t, uint32_t>> rec;
    rec.reserve(events.size());
    for (const Event &e : events) {
        auto it = models.find(e.ctx);
        if (it == models.end()) it = models.emplace(e.ctx, AdaptiveModel(e.alphabet)).first;
        uint32_t l, p;
        it->second.left_and_prob(e.symbol, &l, &p);
        rec.emplace_back(l, p);
        it->second.update(e.symbol);
    }
    return ans_encode(rec);
}

// lib.rs:269-278
Models build_models(const std::vector<Event> &events)
{
    Models models;
    for (const Event &e : events) {
        auto it = models.find(e.ctx);
        if (it == models.end()) it = models.emplace(e.ctx, AdaptiveModel(e.alphabet)).first;
        it->second.update(e.symbol);
    }
    return models;
}

// ---------------------------------------------------------------- format (mars_format.rs)

constexpr uint8_t FIELD_SPLIT = 0, FIELD_MODE = 1, FIELD_QALFA = 2, FIELD_QBETA = 3,
                  FIELD_QBETA_DC = 4, FIELD_ISOMETRY = 5, FIELD_DOM_ROW = 6, FIELD_DOM_COL = 7,
                  FIELD_GX = 8, FIELD_GY = 9, FIELD_RESID_NZ = 10, FIELD_RESID_MAG = 11;
constexpr uint32_t MODE_ALPHABET = 4;
constexpr int32_t AFFINE_GRAD_CLAMP = 2048;
constexpr double AFFINE_GRAD_SCALE = 64.0;
constexpr int32_t LEVEL_CLAMP = 63; // residual.rs:6
constexpr uint32_t BANDS = 4;       // residual.rs:8

uint32_t zigzag64(int64_t x) { return (uint32_t)((x << 1) ^ (x >> 63)); }
uint32_t zigzag32(int32_t x) { return (uint32_t)((x << 1) ^ (x >> 31)); }

struct Predictor {
    int64_t prev_row_units = 0, prev_col_units = 0;
};

// residual.rs:36-59
void residual_events(const std::vector<int32_t> &levels, uint32_t size_class,
                     std::vector<Event> &ev)
{
    const size_t n = levels.size();
    for (size_t i = 0; i < n; i++) {
        const uint32_t band = std::min<uint32_t>((uint32_t)((i * BANDS) / n), BANDS - 1);
        const int32_t cl = std::max(-LEVEL_CLAMP, std::min(LEVEL_CLAMP, levels[i]));
        const bool nz = cl != 0;
        ev.push_back({{FIELD_RESID_NZ, size_class * BANDS + band}, 2, nz ? 1u : 0u});
        if (nz)
            ev.push_back({{FIELD_RESID_MAG, size_class * BANDS + band},
                          (uint32_t)(2 * LEVEL_CLAMP + 1), zigzag32(cl)});
    }
}

// mars_format.rs:515-594
void emit_leaf(const Header &h, const Leaf &leaf, uint32_t size_class, Predictor &pred,
               std::vector<Event> &ev)
{
    ev.push_back({{FIELD_MODE, size_class}, MODE_ALPHABET, leaf.mode});
    if (leaf.mode == 0) {
        ev.push_back({{FIELD_QBETA_DC, size_class}, 1u << h.bits_beta, leaf.qbeta});
    } else if (leaf.mode == 1) {
        ev.push_back({{FIELD_QBETA_DC, size_class}, 1u << h.bits_beta, leaf.qbeta});
        const uint32_t ga = (uint32_t)(2 * (int64_t)AFFINE_GRAD_CLAMP + 1);
        ev.push_back({{FIELD_GX, size_class}, ga, zigzag64(leaf.qgx)});
        ev.push_back({{FIELD_GY, size_class}, ga, zigzag64(leaf.qgy)});
    } else {
        ev.push_back({{FIELD_QALFA, size_class}, (1u << h.bits_alfa) - 1, leaf.qalfa - 1});
        ev.push_back({{FIELD_QBETA, size_class}, 1u << h.bits_beta, leaf.qbeta});
        ev.push_back({{FIELD_ISOMETRY, size_class}, 8, leaf.isometry});
        const int64_t ru = leaf.dom_row / h.shift, cu = leaf.dom_col / h.shift;
        const int64_t rr = 1ll << h.bits_coord_row(), cr = 1ll << h.bits_coord_col();
        ev.push_back({{FIELD_DOM_ROW, size_class}, (uint32_t)(2 * rr),
                      zigzag64(ru - pred.prev_row_units)});
        ev.push_back({{FIELD_DOM_COL, size_class}, (uint32_t)(2 * cr),
                      zigzag64(cu - pred.prev_col_units)});
        pred.prev_row_units = ru;
        pred.prev_col_units = cu;
        if (leaf.mode == 3) residual_events(leaf.residual, size_class, ev);
    }
}

std::vector<Event> leaf_events(const Header &h, const Leaf &leaf, uint32_t size_class)
{
    std::vector<Event> ev;
    Predictor pred;
    emit_leaf(h, leaf, size_class, pred, ev);
    return ev;
}

typedef std::unordered_map<uint64_t, const Leaf *> ByPos;
uint64_t pos_key(uint32_t row, uint32_t col, uint32_t size)
{
    return ((uint64_t)row << 40) | ((uint64_t)col << 16) | size;
}

// mars_format.rs:461-504; returns false on a missing leaf.
bool walk_write(const Header &h, uint32_t row, uint32_t col, uint32_t size, const ByPos &by_pos,
                Predictor &pred, std::vector<Event> &ev)
{
    if (row >= h.height || col >= h.width) return true;
    const uint32_t half = size / 2;
    const bool forced = size > h.max_size || row + size > h.height || col + size > h.width;
    const uint32_t sc = (uint32_t)__builtin_ctz(size);
    auto it = by_pos.find(pos_key(row, col, size));
    if (it != by_pos.end()) {
        if (!forced && size > h.min_size) ev.push_back({{FIELD_SPLIT, sc}, 2, 0});
        emit_leaf(h, *it->second, sc, pred, ev);
        return true;
    }
    if (!forced) {
        if (size <= h.min_size) return false;
        ev.push_back({{FIELD_SPLIT, sc}, 2, 1});
    }
    return walk_write(h, row, col, half, by_pos, pred, ev) &&
           walk_write(h, row + half, col, half, by_pos, pred, ev) &&
           walk_write(h, row, col + half, half, by_pos, pred, ev) &&
           walk_write(h, row + half, col + half, half, by_pos, pred, ev);
}

int build_events(const Header &h, const std::vector<Leaf> &leaves, std::vector<Event> &ev)
{
    ByPos by_pos;
    by_pos.reserve(leaves.size() * 2);
    for (const Leaf &l : leaves)
        if (!by_pos.emplace(pos_key(l.row, l.col, l.size), &l).second) return MARS_HIP_E_ARG;
    Predictor pred;
    if (!walk_write(h, 0, 0, h.virtual_size(), by_pos, pred, ev)) return MARS_HIP_E_ARG;
    return MARS_HIP_OK;
}

// mars_format.rs:232-297 (the 24-byte header, one TREE section).
int format_write(const Header &h, const std::vector<Leaf> &leaves, std::vector<uint8_t> &out)
{
    const bool valid = h.width != 0 && h.height != 0 && h.shift != 0 && h.bits_alfa >= 2 &&
                       h.bits_alfa <= 24 && h.bits_beta >= 1 && h.bits_beta <= 24 &&
                       h.min_size != 0 && h.max_size >= h.min_size;
    if (!valid) return MARS_HIP_E_ARG;
    const uint64_t vs = h.virtual_size(), ms = h.min_size;
    if ((vs / ms) * (vs / ms) > (1ull << 24)) return MARS_HIP_E_ARG;
    if (h.width > 65535 || h.height > 65535) return MARS_HIP_E_ARG;
    if (h.min_size > 255 || h.max_size > 255 || h.shift > 255) return MARS_HIP_E_ARG;
    std::vector<Event> ev;
    int rc = build_events(h, leaves, ev);
    if (rc) return rc;
    const std::vector<uint8_t> tree = entropy_encode(ev);
    out.clear();
    const uint8_t magic[4] = {'M', 'A', 'R', 'S'};
    out.insert(out.end(), magic, magic + 4);
    out.push_back(0); // version
    out.push_back(0); // flags
    out.push_back(1); // channels
    out.push_back(8); // bit depth
    out.push_back(0); // colour space
    out.push_back((uint8_t)h.min_size);
    out.push_back((uint8_t)h.max_size);
    out.push_back((uint8_t)h.shift);
    out.push_back((uint8_t)h.bits_alfa);
    out.push_back((uint8_t)h.bits_beta);
    out.push_back((uint8_t)h.int_max_alfa);
    out.push_back((uint8_t)h.width);
    out.push_back((uint8_t)(h.width >> 8));
    out.push_back((uint8_t)h.height);
    out.push_back((uint8_t)(h.height >> 8));
    out.push_back(1); // section count
    for (int i = 0; i < 4; i++) out.push_back((uint8_t)(h.qstep_bits >> (8 * i)));
    out.push_back(1); // TREE
    const uint32_t tl = (uint32_t)tree.size();
    for (int i = 0; i < 4; i++) out.push_back((uint8_t)(tl >> (8 * i)));
    out.insert(out.end(), tree.begin(), tree.end());
    return MARS_HIP_OK;
}

// ---------------------------------------------------------------- read side

double dead_zone_dequantize(int32_t level, double step, double dz);
std::vector<double> inverse_dct2d(const std::vector<double> &coef, size_t size);
uint32_t qstep_bits(double step);

// The decode half of the stack rANS coder above (constriction's AnsCoder
// from_compressed / decode_symbol, as the reference's Decoder uses it,
// lib.rs:292-334): bytes become little-endian words (a trailing partial word
// is dropped); the state is read from the end of the words -- the first word
// popped must be nonzero, and a second is shifted in below it while the
// state is below 2^32; each symbol takes quantile = state mod 2^24, then
// state = (state >> 24) * p + (quantile - left), and a word is shifted in
// whenever the state falls below 2^32 and words remain.
struct AnsDecoder {
    std::vector<uint32_t> bulk;
    uint64_t state = 0;
    bool init(const uint8_t *p, size_t n)
    {
        bulk.resize(n / 4);
        for (size_t i = 0; i < bulk.size(); i++)
            bulk[i] = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) |
                      ((uint32_t)p[4 * i + 2] << 16) | ((uint32_t)p[4 * i + 3] << 24);
        state = 0;
        if (bulk.empty()) return true;
        const uint32_t first = bulk.back();
        bulk.pop_back();
        if (first == 0) return false; // never produced by the encoder
        state = first;
        while (!bulk.empty() && state < (1ull << 32)) {
            state = (state << 32) | bulk.back();
            bulk.pop_back();
        }
        return true;
    }
    uint32_t decode(const AdaptiveModel &m)
    {
        const uint32_t quantile = (uint32_t)(state & (PRECISION_TOTAL - 1));
        uint32_t sym, left, prob;
        m.symbol_for_quantile(quantile, &sym, &left, &prob);
        state = (state >> PRECISION) * prob + (quantile - left);
        if (state < (1ull << 32) && !bulk.empty()) {
            state = (state << 32) | bulk.back();
            bulk.pop_back();
        }
        return sym;
    }
};

// The reader's largest alphabet.  The format allows bits_alfa/bits_beta up
// to 24 (mars_format.rs:157), but the reference's AdaptiveModel::new
// asserts an alphabet below 2^24 (so a bits_beta = 24 stream panics there)
// and every decode scans the whole alphabet; this reader refuses alphabets
// above 2^20 (bits above 20) with MARS_FMT_E_UNSUPPORTED instead, which keeps
// the work per symbol bounded on arbitrary input.
constexpr uint32_t kMaxReadAlphabet = 1u << 20;

// lib.rs:312-334 Decoder: one model per context, created on first use with
// that event's alphabet.
struct EntropyDecoder {
    AnsDecoder ans;
    Models models;
    bool bad = false; // an alphabet this reader does not take
    uint32_t next(ContextKey ctx, uint32_t alphabet)
    {
        auto it = models.find(ctx);
        if (it == models.end()) {
            if (alphabet < 2 || alphabet > kMaxReadAlphabet) {
                bad = true;
                return 0;
            }
            it = models.emplace(ctx, AdaptiveModel(alphabet)).first;
        }
        const uint32_t s = ans.decode(it->second);
        it->second.update(s);
        return s;
    }
};

int64_t unzigzag64(uint32_t z) { return (int64_t)(z >> 1) ^ -(int64_t)(z & 1); }

// mars_format.rs:596-718 walk_read.
int walk_read(const Header &h, uint32_t row, uint32_t col, uint32_t size, EntropyDecoder &dec,
              Predictor &pred, std::vector<Leaf> &leaves)
{
    if (row >= h.height || col >= h.width) return MARS_HIP_OK;
    const uint32_t half = size / 2;
    const bool forced = size > h.max_size || row + size > h.height || col + size > h.width;
    const uint32_t sc = (uint32_t)__builtin_ctz(size);
    bool split = forced;
    if (!forced && size > h.min_size) split = dec.next({FIELD_SPLIT, sc}, 2) == 1;
    if (dec.bad) return MARS_FMT_E_UNSUPPORTED;
    if (split) {
        int rc = walk_read(h, row, col, half, dec, pred, leaves);
        if (!rc) rc = walk_read(h, row + half, col, half, dec, pred, leaves);
        if (!rc) rc = walk_read(h, row, col + half, half, dec, pred, leaves);
        if (!rc) rc = walk_read(h, row + half, col + half, half, dec, pred, leaves);
        return rc;
    }
    Leaf leaf{};
    leaf.row = row;
    leaf.col = col;
    leaf.size = size;
    const uint32_t mode = dec.next({FIELD_MODE, sc}, MODE_ALPHABET);
    leaf.mode = (uint8_t)mode;
    if (mode <= 1) {
        leaf.qbeta = dec.next({FIELD_QBETA_DC, sc}, 1u << h.bits_beta);
        if (mode == 1) {
            const uint32_t ga = (uint32_t)(2 * (int64_t)AFFINE_GRAD_CLAMP + 1);
            leaf.qgx = (int32_t)unzigzag64(dec.next({FIELD_GX, sc}, ga));
            leaf.qgy = (int32_t)unzigzag64(dec.next({FIELD_GY, sc}, ga));
        }
        if (dec.bad) return MARS_FMT_E_UNSUPPORTED;
        leaves.push_back(std::move(leaf));
        return MARS_HIP_OK;
    }
    leaf.qalfa = dec.next({FIELD_QALFA, sc}, (1u << h.bits_alfa) - 1) + 1;
    leaf.qbeta = dec.next({FIELD_QBETA, sc}, 1u << h.bits_beta);
    leaf.isometry = (uint8_t)dec.next({FIELD_ISOMETRY, sc}, 8);
    const int64_t rr = 1ll << h.bits_coord_row(), cr = 1ll << h.bits_coord_col();
    const int64_t rd = unzigzag64(dec.next({FIELD_DOM_ROW, sc}, (uint32_t)(2 * rr)));
    const int64_t cd = unzigzag64(dec.next({FIELD_DOM_COL, sc}, (uint32_t)(2 * cr)));
    if (dec.bad) return MARS_FMT_E_UNSUPPORTED;
    const int64_t ru = pred.prev_row_units + rd, cu = pred.prev_col_units + cd;
    pred.prev_row_units = ru;
    pred.prev_col_units = cu;
    // (In 64 bits: the reference multiplies in u32, which a corrupt stream's
    // accumulated deltas can overflow.)
    if (ru < 0 || cu < 0 || ru > 0xffffffffll || cu > 0xffffffffll)
        return MARS_FMT_E_DOMAIN_OUT_OF_BOUNDS;
    const uint64_t dr = (uint64_t)h.shift * (uint64_t)ru, dc = (uint64_t)h.shift * (uint64_t)cu;
    if (dr + 2ull * size > h.height || dc + 2ull * size > h.width) return MARS_FMT_E_DOMAIN_OUT_OF_BOUNDS;
    leaf.dom_row = (uint32_t)dr;
    leaf.dom_col = (uint32_t)dc;
    if (mode == 3) {
        // residual.rs:64-80 decode_values
        const size_t n = (size_t)size * size;
        leaf.residual.resize(n);
        for (size_t i = 0; i < n; i++) {
            const uint32_t band = std::min<uint32_t>((uint32_t)((i * BANDS) / n), BANDS - 1);
            if (dec.next({FIELD_RESID_NZ, sc * BANDS + band}, 2) == 1) {
                const uint32_t z = dec.next({FIELD_RESID_MAG, sc * BANDS + band},
                                            (uint32_t)(2 * LEVEL_CLAMP + 1));
                leaf.residual[i] = (int32_t)(z >> 1) ^ -(int32_t)(z & 1);
            }
        }
    }
    leaves.push_back(std::move(leaf));
    return MARS_HIP_OK;
}

bool header_fields_valid(const Header &h)
{
    return h.width != 0 && h.height != 0 && h.shift != 0 && h.bits_alfa >= 2 && h.bits_alfa <= 24 &&
           h.bits_beta >= 1 && h.bits_beta <= 24 && h.min_size != 0 && h.max_size >= h.min_size;
}

bool leaf_count_within_bound(const Header &h)
{
    const uint64_t vs = h.virtual_size(), ms = h.min_size;
    return (vs / ms) * (vs / ms) <= (1ull << 24);
}

// mars_format.rs:302-388 read.
int format_read(const uint8_t *data, size_t len, Header &h, std::vector<Leaf> &leaves)
{
    leaves.clear();
    if (len < 24) return MARS_FMT_E_TRUNCATED;
    if (memcmp(data, "MARS", 4) != 0) return MARS_FMT_E_BAD_MAGIC;
    if (data[4] != 0) return MARS_FMT_E_UNSUPPORTED_VERSION;
    uint32_t qb = (uint32_t)data[20] | ((uint32_t)data[21] << 8) | ((uint32_t)data[22] << 16) |
                  ((uint32_t)data[23] << 24);
    float qf;
    memcpy(&qf, &qb, 4);
    const double qd = (double)qf;
    if (!(std::isfinite(qd) && qd >= 1.0 && qd <= 65535.0)) return MARS_FMT_E_INVALID_QSTEP;
    h = Header{};
    h.qstep_bits = qb;
    h.min_size = data[9];
    h.max_size = data[10];
    h.shift = data[11];
    h.bits_alfa = data[12];
    h.bits_beta = data[13];
    h.int_max_alfa = data[14];
    h.width = (uint32_t)data[15] | ((uint32_t)data[16] << 8);
    h.height = (uint32_t)data[17] | ((uint32_t)data[18] << 8);
    const uint32_t nsec = data[19];
    if (!header_fields_valid(h) || !leaf_count_within_bound(h)) return MARS_FMT_E_DEGENERATE_HEADER;
    size_t off = 24;
    const uint8_t *tree = nullptr;
    size_t tree_len = 0;
    bool have_tree = false;
    for (uint32_t s = 0; s < nsec; s++) {
        if (off + 5 > len) return MARS_FMT_E_TRUNCATED;
        const uint8_t id = data[off];
        const size_t sl = (size_t)data[off + 1] | ((size_t)data[off + 2] << 8) |
                          ((size_t)data[off + 3] << 16) | ((size_t)data[off + 4] << 24);
        off += 5;
        if (sl > len - off) return MARS_FMT_E_SECTION_OUT_OF_BOUNDS;
        if (id == 1) { // the last TREE section wins, as in the reference
            tree = data + off;
            tree_len = sl;
            have_tree = true;
        }
        off += sl;
    }
    if (!have_tree) return MARS_FMT_E_MISSING_TREE;
    EntropyDecoder dec;
    if (!dec.ans.init(tree, tree_len)) return MARS_FMT_E_INVALID_TREE_PAYLOAD;
    Predictor pred;
    const int rc = walk_read(h, 0, 0, h.virtual_size(), dec, pred, leaves);
    if (rc) leaves.clear();
    return rc;
}

// ifs.rs:120-278: the legacy raw `.ifs` bitstream (Mars 1): bits leave each
// byte most significant first, values are assembled least significant bit
// first; header 4+4+7+7+6+12+12+8 bits; the tree NW, SW, NE, SE; 0-7 zero
// padding bits.
struct BitReader {
    const uint8_t *data;
    size_t len, pos = 0;
    bool ok = true;
    uint32_t read(uint32_t n)
    {
        uint32_t v = 0;
        for (uint32_t i = 0; i < n; i++) {
            if ((pos >> 3) >= len) {
                ok = false;
                return 0;
            }
            const uint32_t b = (data[pos >> 3] >> (7 - (pos & 7))) & 1u;
            pos++;
            v |= b << i;
        }
        return v;
    }
};

int ifs_walk(BitReader &r, const Header &h, uint32_t row, uint32_t col, uint32_t size,
             std::vector<Leaf> &leaves)
{
    if (row >= h.height || col >= h.width) return MARS_HIP_OK;
    const bool forced = size > h.max_size || row + size > h.height || col + size > h.width;
    bool split = forced;
    if (!forced && size > h.min_size) {
        split = r.read(1) == 1;
        if (!r.ok) return MARS_FMT_E_TRUNCATED;
    }
    if (split) {
        const uint32_t half = size / 2;
        int rc = ifs_walk(r, h, row, col, half, leaves);
        if (!rc) rc = ifs_walk(r, h, row + half, col, half, leaves);
        if (!rc) rc = ifs_walk(r, h, row, col + half, half, leaves);
        if (!rc) rc = ifs_walk(r, h, row + half, col + half, half, leaves);
        return rc;
    }
    Leaf l{};
    l.row = row;
    l.col = col;
    l.size = size;
    l.qalfa = r.read(h.bits_alfa);
    l.qbeta = r.read(h.bits_beta);
    if (l.qalfa != 0) {
        l.isometry = (uint8_t)r.read(3);
        const uint64_t dr = (uint64_t)h.shift * r.read(h.bits_coord_row());
        const uint64_t dc = (uint64_t)h.shift * r.read(h.bits_coord_col());
        if (!r.ok) return MARS_FMT_E_TRUNCATED;
        if (dr + 2ull * size > h.height || dc + 2ull * size > h.width)
            return MARS_FMT_E_DOMAIN_OUT_OF_BOUNDS;
        l.dom_row = (uint32_t)dr;
        l.dom_col = (uint32_t)dc;
    }
    if (!r.ok) return MARS_FMT_E_TRUNCATED;
    l.mode = l.qalfa == 0 ? 0 : 2;
    leaves.push_back(std::move(l));
    return MARS_HIP_OK;
}

int ifs_parse(const uint8_t *data, size_t len, Header &h, std::vector<Leaf> &leaves)
{
    leaves.clear();
    BitReader r{data, len};
    h = Header{};
    h.bits_alfa = r.read(4);
    h.bits_beta = r.read(4);
    h.min_size = r.read(7);
    h.max_size = r.read(7);
    h.shift = r.read(6);
    h.width = r.read(12);
    h.height = r.read(12);
    h.int_max_alfa = r.read(8);
    h.qstep_bits = qstep_bits(8.0); // ResidualQstep::LEGACY
    if (!r.ok) return MARS_FMT_E_TRUNCATED;
    if (h.width == 0 || h.height == 0 || h.shift == 0) return MARS_FMT_E_DEGENERATE_HEADER;
    // (min_size 0 or max_size < min_size would split forever: the reference
    // recurses until the bitstream runs out; here it is a degenerate header.)
    if (h.min_size == 0 || h.max_size < h.min_size) return MARS_FMT_E_DEGENERATE_HEADER;
    int rc = ifs_walk(r, h, 0, 0, h.virtual_size(), leaves);
    if (rc) {
        leaves.clear();
        return rc;
    }
    const size_t left = len * 8 - r.pos;
    if (left >= 8) rc = MARS_FMT_E_TRAILING_BITS;
    else if (left > 0 && (data[len - 1] & ((1u << left) - 1)) != 0) rc = MARS_FMT_E_NONZERO_PADDING;
    if (rc) leaves.clear();
    return rc;
}

// ifs.rs:406-484 decode_leaf: one leaf of the next iteration from the previous
// one (double-buffered, so leaves are independent).
void decode_leaf(const Header &h, const Leaf &leaf, const uint8_t *img, size_t stride,
                 uint8_t *next)
{
    const size_t size = leaf.size;
    auto to_u8 = [](double v) -> uint8_t {
        if (!(v == v)) return 0; // Rust's `as u8` of NaN
        v = std::min(std::max(v, 0.0), 255.0);
        return (uint8_t)v;
    };
    if (leaf.mode == 1) {
        const double b0 = (double)leaf.qbeta / (double)((1u << h.bits_beta) - 1) * 255.0;
        const double gx = (double)leaf.qgx / AFFINE_GRAD_SCALE;
        const double gy = (double)leaf.qgy / AFFINE_GRAD_SCALE;
        for (size_t u = 0; u < size; u++)
            for (size_t v = 0; v < size; v++)
                next[(leaf.row + u) * stride + leaf.col + v] =
                    to_u8(std::round(b0 + gx * (double)u + gy * (double)v));
        return;
    }
    const double max_alfa = (double)h.int_max_alfa / 256.0 * 8.0;
    const double alfa = (double)leaf.qalfa / (double)(1u << h.bits_alfa) * max_alfa;
    double beta = (double)leaf.qbeta / (double)((1u << h.bits_beta) - 1) * ((1.0 + std::fabs(alfa)) * 255.0);
    if (alfa > 0.0) beta -= alfa * 255.0;
    std::vector<double> res;
    if (leaf.mode == 3 && !leaf.residual.empty()) {
        std::vector<double> levels(leaf.residual.size());
        for (size_t k = 0; k < levels.size(); k++)
            levels[k] = dead_zone_dequantize(leaf.residual[k], h.qstep(), 0.5);
        res = inverse_dct2d(levels, size);
    }
    for (size_t u = 0; u < size; u++)
        for (size_t v = 0; v < size; v++) {
            double d = 0.0;
            if (leaf.qalfa != 0) {
                const size_t dr = leaf.dom_row + 2 * u, dc = leaf.dom_col + 2 * v;
                d = ((double)img[dr * stride + dc] + (double)img[(dr + 1) * stride + dc] +
                     (double)img[dr * stride + dc + 1] + (double)img[(dr + 1) * stride + dc + 1]) /
                    4.0;
            }
            uint32_t i, j;
            iso_map(leaf.isometry, (uint32_t)u, (uint32_t)v, (uint32_t)size, &i, &j);
            double value = (0.5 + d * alfa) + beta;
            if (!res.empty()) value += res[u * size + v];
            next[(leaf.row + i) * stride + leaf.col + j] = to_u8(value);
        }
}

// ---------------------------------------------------------------- rate (rate.rs:60-85)

struct RateModels {
    Models models;
    double bits_for(ContextKey ctx, uint32_t alphabet, uint32_t symbol) const
    {
        auto it = models.find(ctx);
        if (it != models.end()) return it->second.bits_for(symbol);
        return AdaptiveModel(std::max<uint32_t>(alphabet, 2)).bits_for(symbol);
    }
};

// encode.rs:1515-1521 event_bits: f64 sum in event order.
double event_bits(const RateModels &rate, const std::vector<Event> &ev)
{
    double s = -0.0;
    for (const Event &e : ev) s += rate.bits_for(e.ctx, e.alphabet, e.symbol);
    return s;
}

// ---------------------------------------------------------------- quantisers

// encode.rs:446-448
uint32_t quantise_f64(double x, uint32_t max)
{
    double t = std::trunc(0.5 + x);
    if (t < 0.0) t = 0.0;
    if (t > (double)max) t = (double)max;
    return (uint32_t)t;
}

// encode.rs:456-460
uint32_t best_beta(int64_t range_sum, int64_t pixel_count, uint32_t bits_beta)
{
    const double mean = (double)range_sum / (double)pixel_count;
    const uint32_t mq = (1u << bits_beta) - 1;
    return quantise_f64(mean / 255.0 * (double)mq, mq);
}

// quant.rs:71-82
int32_t dead_zone_quantize(double x, double step, double dz)
{
    const double threshold = dz * step;
    const double ax = std::fabs(x);
    if (ax <= threshold) return 0;
    const double level = std::floor((ax - threshold) / step) + 1.0;
    const double signed_level = x < 0.0 ? -level : level;
    return (int32_t)signed_level;
}

// quant.rs:87-98
double dead_zone_dequantize(int32_t level, double step, double dz)
{
    if (level == 0) return 0.0;
    const double threshold = dz * step;
    const double mag = threshold + ((double)(uint32_t)std::abs(level) - 1.0 + 0.5) * step;
    return level < 0 ? -mag : mag;
}

// quant.rs:28-50 ResidualQstep::new / from_lambda: binary32 bits.
uint32_t qstep_bits(double step)
{
    const float f = (float)step;
    uint32_t b;
    memcpy(&b, &f, 4);
    return b;
}
uint32_t qstep_from_lambda(double lambda)
{
    double step = std::sqrt(std::max(lambda, 0.0)) * std::sqrt(6.0 / 0.6931471805599453);
    step = std::min(std::max(step, 1.0), 65535.0);
    return qstep_bits(step);
}

// ---------------------------------------------------------------- DCT (dct.rs)

double dct_basis(size_t N, size_t k, size_t n)
{
    const double ck = k == 0 ? 0.7071067811865476 : 1.0;
    return std::sqrt(2.0 / (double)N) * ck *
           std::cos(3.141592653589793 / (double)N * ((double)n + 0.5) * (double)k);
}

void dct1d(const double *x, size_t n, double *out, bool inverse)
{
    for (size_t a = 0; a < n; a++) {
        double s = -0.0;
        for (size_t b = 0; b < n; b++)
            s += x[b] * (inverse ? dct_basis(n, b, a) : dct_basis(n, a, b));
        out[a] = s;
    }
}

std::vector<double> forward_dct2d(const std::vector<double> &blk, size_t size)
{
    std::vector<double> tmp(size * size), out(size * size), col(size), t(size);
    for (size_t r = 0; r < size; r++) dct1d(&blk[r * size], size, &tmp[r * size], false);
    for (size_t c = 0; c < size; c++) {
        for (size_t r = 0; r < size; r++) col[r] = tmp[r * size + c];
        dct1d(col.data(), size, t.data(), false);
        for (size_t r = 0; r < size; r++) out[r * size + c] = t[r];
    }
    return out;
}

std::vector<double> inverse_dct2d(const std::vector<double> &coef, size_t size)
{
    std::vector<double> tmp(size * size), out(size * size), col(size), t(size);
    for (size_t c = 0; c < size; c++) {
        for (size_t r = 0; r < size; r++) col[r] = coef[r * size + c];
        dct1d(col.data(), size, t.data(), true);
        for (size_t r = 0; r < size; r++) tmp[r * size + c] = t[r];
    }
    for (size_t r = 0; r < size; r++) dct1d(&tmp[r * size], size, &out[r * size], true);
    return out;
}

// ---------------------------------------------------------------- encode (encode.rs)

struct Params {
    uint32_t min_size, max_size, shift, bits_alfa, bits_beta, zero_threshold;
    double max_alfa, t_rms;
    bool has_lambda;
    double lambda;
};

struct Provider {
    mars_search_fn fn;
    void *user;
};

struct Ctx {
    const uint8_t *px;
    uint32_t w, h;
    Provider provider;
    const Header *hdr;
    const Params *params;
    const RateModels *rate;
    bool allowed[4];
    bool adaptive_density;
    uint32_t rd_candidates; // P5c: ranked candidates the RD walk asks for (<= 1: the winner)
    int err;
};

// The provider's answer: SearchOutcome (n <= 1) or RankedCandidates
// (encode.rs:536-555), ascending by rms; c[0] is the winner.
struct Found {
    bool ok;
    uint32_t n;
    mars_hip_candidate c[MARS_MAX_RANKED];
    uint64_t evals;
};

Found do_search(Ctx &ctx, uint32_t row, uint32_t col, uint32_t size, uint32_t shift,
                uint32_t max = 1)
{
    Found f{};
    uint32_t n = 0;
    if (max < 1) max = 1;
    const int rc =
        ctx.provider.fn(ctx.provider.user, row, col, size, shift, max, f.c, &n, &f.evals);
    if (rc && !ctx.err) ctx.err = rc;
    f.ok = rc == 0 && n > 0;
    f.n = f.ok ? std::min(n, max) : 0;
    return f;
}

int64_t block_sum(const Ctx &ctx, uint32_t row, uint32_t col, uint32_t size)
{
    int64_t s = 0;
    for (uint32_t i = 0; i < size; i++)
        for (uint32_t j = 0; j < size; j++) s += ctx.px[(size_t)(row + i) * ctx.w + col + j];
    return s;
}

// encode.rs:1330-1435 walk (threshold) + 1442-1468 split.
void walk(Ctx &ctx, uint32_t row, uint32_t col, uint32_t size, std::vector<Leaf> &leaves,
          uint64_t &evals)
{
    const Header &h = *ctx.hdr;
    if (row >= h.height || col >= h.width) return;
    const bool forced = size > h.max_size || row + size > h.height || col + size > h.width;
    if (forced) {
        const uint32_t half = size / 2;
        walk(ctx, row, col, half, leaves, evals);
        walk(ctx, row + half, col, half, leaves, evals);
        walk(ctx, row, col + half, half, leaves, evals);
        walk(ctx, row + half, col + half, half, leaves, evals);
        return;
    }
    if (size == 1) {
        const uint32_t pixel = ctx.px[(size_t)row * ctx.w + col];
        leaves.push_back(dc_leaf(row, col, size, best_beta(pixel, 1, h.bits_beta)));
        return;
    }
    const Found f = do_search(ctx, row, col, size, ctx.params->shift);
    const double best_rms = f.ok ? f.c[0].rms : INFINITY;
    if (best_rms > ctx.params->t_rms && size > h.min_size) {
        const uint32_t half = size / 2;
        walk(ctx, row, col, half, leaves, evals);
        walk(ctx, row + half, col, half, leaves, evals);
        walk(ctx, row, col + half, half, leaves, evals);
        walk(ctx, row + half, col + half, half, leaves, evals);
        evals += f.evals;
        return;
    }
    Leaf leaf = dc_leaf(row, col, size, 0);
    if (f.ok) {
        leaf.mode = 2;
        leaf.qalfa = f.c[0].qalfa;
        leaf.qbeta = f.c[0].qbeta;
        leaf.isometry = (uint8_t)f.c[0].isometry;
        leaf.dom_row = f.c[0].dom_row;
        leaf.dom_col = f.c[0].dom_col;
    }
    if (leaf.qalfa <= ctx.params->zero_threshold) { // section 8.1 zero-alfa override
        leaf.qbeta = best_beta(block_sum(ctx, row, col, size), (int64_t)size * size, h.bits_beta);
        leaf.qalfa = 0;
        leaf.mode = 0;
        leaf.isometry = 0;
        leaf.dom_row = 0;
        leaf.dom_col = 0;
    }
    leaves.push_back(std::move(leaf));
    evals += f.evals;
}

// encode.rs:1721-1789 affine_fit: returns (qbeta, qgx, qgy, sse).
void affine_fit(const std::vector<double> &block, size_t size, uint32_t bits_beta, uint32_t *qb,
                int32_t *qgx, int32_t *qgy, double *sse_out)
{
    const double n = (double)size * (double)size;
    double t0 = 0.0, tu = 0.0, tv = 0.0;
    for (size_t u = 0; u < size; u++)
        for (size_t v = 0; v < size; v++) {
            const double p = block[u * size + v];
            t0 += p;
            tu += p * (double)u;
            tv += p * (double)v;
        }
    double su = -0.0, suu = -0.0;
    for (size_t u = 0; u < size; u++) su += (double)u;
    for (size_t u = 0; u < size; u++) suu += (double)u * (double)u;
    const double s0 = n, s_u = (double)size * su, s_v = s_u, s_uu = (double)size * suu,
                 s_vv = s_uu, s_uv = su * su;
    auto det3 = [](const double m[3][3]) {
        return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
               m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    };
    const double A[3][3] = {{s0, s_u, s_v}, {s_u, s_uu, s_uv}, {s_v, s_uv, s_vv}};
    const double det_a = det3(A);
    double b0, gx, gy;
    if (std::fabs(det_a) < 1e-9) {
        b0 = t0 / s0;
        gx = 0.0;
        gy = 0.0;
    } else {
        const double Ab0[3][3] = {{t0, s_u, s_v}, {tu, s_uu, s_uv}, {tv, s_uv, s_vv}};
        const double Agx[3][3] = {{s0, t0, s_v}, {s_u, tu, s_uv}, {s_v, tv, s_vv}};
        const double Agy[3][3] = {{s0, s_u, t0}, {s_u, s_uu, tu}, {s_v, s_uv, tv}};
        b0 = det3(Ab0) / det_a;
        gx = det3(Agx) / det_a;
        gy = det3(Agy) / det_a;
    }
    const uint32_t mq = (1u << bits_beta) - 1;
    const uint32_t qbeta = quantise_f64(std::max(b0 / 255.0 * (double)mq, 0.0), mq);
    auto qgrad = [](double g) {
        double r = std::round(g * AFFINE_GRAD_SCALE);
        r = std::max(-(double)AFFINE_GRAD_CLAMP, std::min((double)AFFINE_GRAD_CLAMP, r));
        return (int32_t)r;
    };
    const int32_t qx = qgrad(gx), qy = qgrad(gy);
    const double b0q = (double)qbeta / (double)mq * 255.0;
    const double gxq = (double)qx / AFFINE_GRAD_SCALE, gyq = (double)qy / AFFINE_GRAD_SCALE;
    double sse = 0.0;
    for (size_t u = 0; u < size; u++)
        for (size_t v = 0; v < size; v++) {
            const double pred = b0q + gxq * (double)u + gyq * (double)v;
            const double diff = block[u * size + v] - pred;
            sse += diff * diff;
        }
    *qb = qbeta;
    *qgx = qx;
    *qgy = qy;
    *sse_out = sse;
}

// encode.rs:1798-1862 residual_for_candidate: returns (levels, sse).
double residual_for_candidate(const Ctx &ctx, const std::vector<double> &block, size_t size,
                              const mars_hip_candidate &c, std::vector<int32_t> &levels)
{
    const Params &P = *ctx.params;
    const Header &h = *ctx.hdr;
    const double alfa = (double)c.qalfa / (double)(1u << P.bits_alfa) * P.max_alfa;
    double beta = (double)c.qbeta / (double)((1u << h.bits_beta) - 1) *
                  ((1.0 + std::fabs(alfa)) * 255.0);
    if (alfa > 0.0) beta -= alfa * 255.0;
    std::vector<double> resid(size * size), prediction(size * size);
    const uint8_t *px = ctx.px;
    const size_t stride = ctx.w;
    for (size_t u = 0; u < size; u++)
        for (size_t v = 0; v < size; v++) {
            const size_t dr = c.dom_row + 2 * u, dc = c.dom_col + 2 * v;
            const double d = ((double)px[dr * stride + dc] + (double)px[(dr + 1) * stride + dc] +
                              (double)px[dr * stride + dc + 1] +
                              (double)px[(dr + 1) * stride + dc + 1]) /
                             4.0;
            uint32_t i, j;
            iso_map(c.isometry, (uint32_t)u, (uint32_t)v, (uint32_t)size, &i, &j);
            const double pred = 0.5 + d * alfa + beta;
            prediction[u * size + v] = pred;
            resid[u * size + v] = block[i * size + j] - pred;
        }
    const std::vector<double> coeffs = forward_dct2d(resid, size);
    const double step = h.qstep();
    levels.resize(size * size);
    std::vector<double> deq(size * size);
    for (size_t k = 0; k < size * size; k++) {
        levels[k] = std::max(-LEVEL_CLAMP,
                             std::min(LEVEL_CLAMP, dead_zone_quantize(coeffs[k], step, 0.5)));
        deq[k] = dead_zone_dequantize(levels[k], step, 0.5);
    }
    const std::vector<double> rr = inverse_dct2d(deq, size);
    double sse = 0.0;
    for (size_t u = 0; u < size; u++)
        for (size_t v = 0; v < size; v++) {
            const size_t k = u * size + v;
            uint32_t i, j;
            iso_map(c.isometry, (uint32_t)u, (uint32_t)v, (uint32_t)size, &i, &j);
            const double x = std::min(255.0, std::max(0.0, prediction[k] + rr[k]));
            const uint8_t pixel = (uint8_t)x;
            const double e = block[i * size + j] - (double)pixel;
            sse += e * e;
        }
    return sse;
}

struct Scored {
    Leaf leaf;
    double d, r;
};

// encode.rs:1536-1714 best_mode_leaf (ranked = the provider's ascending list:
// the winner, or up to rd_candidates of them).
Scored best_mode_leaf(const Ctx &ctx, uint32_t row, uint32_t col, uint32_t size, const Found &f,
                      double lambda, uint32_t size_class)
{
    const Header &h = *ctx.hdr;
    const RateModels &rate = *ctx.rate;
    const size_t su = size;
    std::vector<double> block(su * su);
    int64_t range_sum = 0;
    for (size_t i = 0; i < su; i++)
        for (size_t j = 0; j < su; j++) {
            const uint8_t v = ctx.px[(row + i) * (size_t)ctx.w + col + j];
            block[i * su + j] = (double)v;
            range_sum += v;
        }
    const int64_t s0 = (int64_t)size * size;
    std::vector<Scored> cands;
    auto mode0 = [&]() {
        const uint32_t qbeta = best_beta(range_sum, s0, h.bits_beta);
        Leaf leaf = dc_leaf(row, col, size, qbeta);
        const uint32_t mq = (1u << h.bits_beta) - 1;
        const double beta2 = (double)qbeta / (double)mq * 255.0;
        double sse = -0.0;
        for (double p : block) sse += (p - beta2) * (p - beta2);
        const double r = event_bits(rate, leaf_events(h, leaf, size_class));
        cands.push_back({std::move(leaf), sse, r});
    };
    if (ctx.allowed[0]) mode0();
    if (ctx.allowed[1]) {
        uint32_t qb;
        int32_t gx, gy;
        double sse;
        affine_fit(block, su, h.bits_beta, &qb, &gx, &gy, &sse);
        Leaf leaf = dc_leaf(row, col, size, qb);
        leaf.mode = 1;
        leaf.qgx = gx;
        leaf.qgy = gy;
        const double r = event_bits(rate, leaf_events(h, leaf, size_class));
        cands.push_back({std::move(leaf), sse, r});
    }
    // encode.rs:1627-1674: modes 2 and 3 for every ranked candidate, in order,
    // skipping those whose fit collapsed to qalfa 0.
    for (uint32_t ci = 0; ci < f.n; ci++) {
        const mars_hip_candidate &c = f.c[ci];
        if (c.qalfa < 1) continue;
        if (ctx.allowed[2]) {
            Leaf leaf = dc_leaf(row, col, size, c.qbeta);
            leaf.mode = 2;
            leaf.qalfa = c.qalfa;
            leaf.isometry = (uint8_t)c.isometry;
            leaf.dom_row = c.dom_row;
            leaf.dom_col = c.dom_col;
            const double sse2 = c.rms * c.rms * (double)size * (double)size;
            const double r = event_bits(rate, leaf_events(h, leaf, size_class));
            cands.push_back({std::move(leaf), sse2, r});
        }
        if (ctx.allowed[3]) {
            Leaf leaf = dc_leaf(row, col, size, c.qbeta);
            leaf.mode = 3;
            leaf.qalfa = c.qalfa;
            leaf.isometry = (uint8_t)c.isometry;
            leaf.dom_row = c.dom_row;
            leaf.dom_col = c.dom_col;
            const double sse3 = residual_for_candidate(ctx, block, su, c, leaf.residual);
            const double r = event_bits(rate, leaf_events(h, leaf, size_class));
            cands.push_back({std::move(leaf), sse3, r});
        }
    }
    if (cands.empty()) mode0();
    // Iterator::min_by: the first of equal minima.
    size_t best = 0;
    for (size_t i = 1; i < cands.size(); i++) {
        const double ji = cands[i].d + lambda * cands[i].r;
        const double jb = cands[best].d + lambda * cands[best].r;
        if (ji < jb) best = i;
    }
    return std::move(cands[best]);
}

struct RdResult {
    std::vector<Leaf> leaves;
    uint64_t evals = 0;
    double d = 0.0, r = 0.0;
    uint64_t leaf_modes[4] = {0, 0, 0, 0};
    uint64_t split_decisions = 0, leaf_decisions = 0;
};

// encode.rs:372-437 block_rms / adaptive_shift.
double block_rms(const Ctx &ctx, uint32_t row, uint32_t col, uint32_t size)
{
    int64_t t0 = 0, t2 = 0;
    for (uint32_t i = 0; i < size; i++)
        for (uint32_t j = 0; j < size; j++) {
            const int64_t v = ctx.px[(size_t)(row + i) * ctx.w + col + j];
            t0 += v;
            t2 += v * v;
        }
    const int64_t s0 = (int64_t)size * size;
    const double mean = (double)t0 / (double)s0;
    const double var = std::max((double)t2 / (double)s0 - mean * mean, 0.0);
    return std::sqrt(var);
}

RdResult walk_rd(Ctx &ctx, uint32_t row, uint32_t col, uint32_t size, double lambda);

// encode.rs:2033-2071 split_rd: the four children in TL, BL, TR, BR order.
RdResult split_rd(Ctx &ctx, uint32_t row, uint32_t col, uint32_t half, double lambda)
{
    RdResult q[4] = {walk_rd(ctx, row, col, half, lambda),
                     walk_rd(ctx, row + half, col, half, lambda),
                     walk_rd(ctx, row, col + half, half, lambda),
                     walk_rd(ctx, row + half, col + half, half, lambda)};
    RdResult out;
    for (RdResult &c : q) {
        for (Leaf &l : c.leaves) out.leaves.push_back(std::move(l));
        out.evals += c.evals;
        out.d += c.d;
        out.r += c.r;
        for (int i = 0; i < 4; i++) out.leaf_modes[i] += c.leaf_modes[i];
        out.split_decisions += c.split_decisions;
        out.leaf_decisions += c.leaf_decisions;
    }
    return out;
}

// encode.rs:1865-2027 walk_rd (no lambda regions: node lambda = lambda).
RdResult walk_rd(Ctx &ctx, uint32_t row, uint32_t col, uint32_t size, double lambda)
{
    const Header &h = *ctx.hdr;
    if (row >= h.height || col >= h.width) return RdResult{};
    const bool forced = size > h.max_size || row + size > h.height || col + size > h.width;
    if (forced) return split_rd(ctx, row, col, size / 2, lambda);
    const RateModels &rate = *ctx.rate;
    const uint32_t size_class = (uint32_t)__builtin_ctz(size);
    if (size == 1) {
        const uint32_t pixel = ctx.px[(size_t)row * ctx.w + col];
        Leaf leaf = dc_leaf(row, col, size, best_beta(pixel, 1, h.bits_beta));
        double x = 0.5 + (double)leaf.qbeta / (double)((1u << h.bits_beta) - 1) * 255.0;
        x = std::min(255.0, std::max(0.0, x));
        const uint8_t decoded = (uint8_t)x;
        const double e = (double)pixel - (double)decoded;
        RdResult out;
        out.d = e * e;
        out.r = event_bits(rate, leaf_events(h, leaf, size_class));
        out.leaf_modes[0] = 1;
        out.leaves.push_back(std::move(leaf));
        return out;
    }
    const double node_lambda = lambda * 1.0;
    uint32_t shift = ctx.params->shift;
    if (ctx.adaptive_density && block_rms(ctx, row, col, size) <= 4.0) shift *= 2;
    // encode.rs:1946-1975: the winner, or the provider's ranked list (P5c).
    const Found f = do_search(ctx, row, col, size, shift, ctx.rd_candidates <= 1 ? 1 : ctx.rd_candidates);
    Scored leaf = best_mode_leaf(ctx, row, col, size, f, node_lambda, size_class);
    if (size <= h.min_size) {
        RdResult out;
        out.evals = f.evals;
        out.d = leaf.d;
        out.r = leaf.r;
        out.leaf_modes[leaf.leaf.mode] = 1;
        out.leaves.push_back(std::move(leaf.leaf));
        return out;
    }
    RdResult children = split_rd(ctx, row, col, size / 2, lambda);
    const uint64_t total_evals = f.evals + children.evals;
    const double b0 = rate.bits_for({FIELD_SPLIT, size_class}, 2, 0);
    const double b1 = rate.bits_for({FIELD_SPLIT, size_class}, 2, 1);
    const double leaf_j = leaf.d + node_lambda * (leaf.r + b0);
    const double split_j = children.d + node_lambda * (children.r + b1);
    if (leaf_j <= split_j) {
        RdResult out;
        out.evals = total_evals;
        out.d = leaf.d;
        out.r = leaf.r + b0;
        out.leaf_modes[leaf.leaf.mode] = 1;
        out.leaf_decisions = 1;
        out.leaves.push_back(std::move(leaf.leaf));
        return out;
    }
    children.evals = total_evals;
    children.r = children.r + b1;
    children.split_decisions += 1;
    return children;
}

} // namespace

extern "C" {

int mars_encode_plane(const uint8_t *px, uint32_t w, uint32_t h, const mars_encode_params *p,
                      const mars_encode_options *o, mars_search_fn search, void *search_user,
                      mars_search_fn warmup, void *warmup_user, uint8_t **out, size_t *out_len,
                      mars_encode_stats *stats)
{
    if (!px || !p || !o || !search || !out || !out_len || w == 0 || h == 0)
        return MARS_HIP_E_ARG;
    if (o->rd_candidates > MARS_MAX_RANKED) return MARS_HIP_E_UNSUPPORTED;
    if (p->has_lambda && !(std::isfinite(p->lambda) && p->lambda >= 0.0)) return MARS_HIP_E_ARG;
    if (p->has_lambda && !warmup) return MARS_HIP_E_ARG;
    *out = nullptr;
    *out_len = 0;
    Params P{p->min_size, p->max_size,       p->shift, p->bits_alfa, p->bits_beta,
             p->zero_threshold, p->max_alfa, p->t_rms, p->has_lambda != 0, p->lambda};
    // encode.rs:1069-1075: the header's residual step.
    uint32_t qb;
    if (o->lambda_adaptive_qstep)
        qb = P.has_lambda ? qstep_from_lambda(P.lambda) : qstep_bits(8.0);
    else {
        if (!(std::isfinite(o->fixed_qstep) && o->fixed_qstep >= 1.0 && o->fixed_qstep <= 65535.0))
            return MARS_HIP_E_ARG;
        qb = qstep_bits(o->fixed_qstep);
    }
    // encode.rs:1033-1047 build_header.
    Header hdr{P.bits_alfa, P.bits_beta, P.min_size, P.max_size, P.shift, w, h,
               quantise_f64(P.max_alfa / 8.0 * 256.0, 255), qb};

    Ctx ctx{};
    ctx.px = px;
    ctx.w = w;
    ctx.h = h;
    ctx.hdr = &hdr;
    ctx.params = &P;
    mars_encode_stats st{};
    std::vector<Leaf> leaves;
    if (P.has_lambda) {
        // encode.rs:1139-1169 build_rate_snapshot: threshold walk at t_rms 8.
        Params wp = P;
        wp.t_rms = MARS_WARMUP_T_RMS;
        wp.has_lambda = false;
        Ctx wctx = ctx;
        wctx.params = &wp;
        wctx.provider = {warmup, warmup_user};
        wctx.rd_candidates = 1; // encode.rs:1160
        wctx.allowed[0] = wctx.allowed[1] = wctx.allowed[2] = wctx.allowed[3] = true;
        std::vector<Leaf> wl;
        uint64_t we = 0;
        walk(wctx, 0, 0, hdr.virtual_size(), wl, we);
        if (wctx.err) return wctx.err;
        std::vector<Event> ev;
        Header gh = hdr;
        gh.qstep_bits = qstep_bits(8.0);
        int rc = build_events(gh, wl, ev);
        if (rc) return rc;
        RateModels rate{build_models(ev)};
        st.warmup_evals = we;

        ctx.provider = {search, search_user};
        ctx.rate = &rate;
        ctx.rd_candidates = o->rd_candidates;
        for (int i = 0; i < 4; i++) ctx.allowed[i] = o->allowed_modes[i] != 0;
        ctx.adaptive_density = o->adaptive_density != 0;
        RdResult res = walk_rd(ctx, 0, 0, hdr.virtual_size(), P.lambda);
        if (ctx.err) return ctx.err;
        leaves = std::move(res.leaves);
        st.search_evals = res.evals;
        for (int i = 0; i < 4; i++) st.leaf_modes[i] = res.leaf_modes[i];
        st.split_decisions = res.split_decisions;
        st.leaf_decisions = res.leaf_decisions;
    } else {
        ctx.provider = {search, search_user};
        for (int i = 0; i < 4; i++) ctx.allowed[i] = o->allowed_modes[i] != 0;
        uint64_t ev = 0;
        walk(ctx, 0, 0, hdr.virtual_size(), leaves, ev);
        if (ctx.err) return ctx.err;
        st.search_evals = ev;
    }
    st.n_leaves = leaves.size();
    std::vector<uint8_t> bytes;
    int rc = format_write(hdr, leaves, bytes);
    if (rc) return rc;
    uint8_t *buf = (uint8_t *)malloc(bytes.size() ? bytes.size() : 1);
    if (!buf) return MARS_HIP_E_ARG;
    memcpy(buf, bytes.data(), bytes.size());
    *out = buf;
    *out_len = bytes.size();
    if (stats) *stats = st;
    return MARS_HIP_OK;
}

int mars_write_container(uint8_t mode, const uint8_t *const *streams, const size_t *lens,
                         uint32_t n_streams, uint8_t **out, size_t *out_len)
{
    if (!out || !out_len || n_streams > 255 || (n_streams && (!streams || !lens)))
        return MARS_HIP_E_ARG;
    // color.rs:778-790: "MARC", version 0, mode, section count, then per
    // stream a u32 little-endian length and the bytes.
    size_t total = 7;
    for (uint32_t i = 0; i < n_streams; i++) {
        if (lens[i] > 0xffffffffu || (lens[i] && !streams[i])) return MARS_HIP_E_ARG;
        total += 4 + lens[i];
    }
    uint8_t *buf = (uint8_t *)malloc(total);
    if (!buf) return MARS_HIP_E_ARG;
    buf[0] = 'M';
    buf[1] = 'A';
    buf[2] = 'R';
    buf[3] = 'C';
    buf[4] = 0;
    buf[5] = mode;
    buf[6] = (uint8_t)n_streams;
    size_t off = 7;
    for (uint32_t i = 0; i < n_streams; i++) {
        for (int k = 0; k < 4; k++) buf[off + k] = (uint8_t)(lens[i] >> (8 * k));
        off += 4;
        if (lens[i]) memcpy(buf + off, streams[i], lens[i]);
        off += lens[i];
    }
    *out = buf;
    *out_len = total;
    return MARS_HIP_OK;
}

int mars_wrap_gray(const uint8_t *plane, size_t len, uint8_t **out, size_t *out_len)
{
    if (!plane && len) return MARS_HIP_E_ARG;
    // color.rs:382-384: mode 0 (grey), one stream.
    return mars_write_container(0, &plane, &len, 1, out, out_len);
}

int mars_ycbcr(const uint8_t *rgb, size_t n, uint8_t *y, uint8_t *cb, uint8_t *cr)
{
    if (n && (!rgb || !y || !cb || !cr)) return MARS_HIP_E_ARG;
    // metrics.rs:394-410: binary64, evaluated left to right without FMA
    // contraction, round() half away from zero, clamp to 0..255.
    auto u8 = [](double x) {
        x = std::round(x);
        x = x < 0.0 ? 0.0 : x > 255.0 ? 255.0 : x;
        return (uint8_t)x;
    };
    for (size_t i = 0; i < n; i++) {
        const double r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
        y[i] = u8(0.299 * r + 0.587 * g + 0.114 * b);
        cb[i] = u8(128.0 - 0.168736 * r - 0.331264 * g + 0.5 * b);
        cr[i] = u8(128.0 + 0.5 * r - 0.418688 * g - 0.081312 * b);
    }
    return MARS_HIP_OK;
}

int mars_downsample_box(const uint8_t *px, uint32_t w, uint32_t h, uint8_t *out)
{
    if (!px || !out || w == 0 || h == 0) return MARS_HIP_E_ARG;
    // color.rs:60-78: (sum of the 2x2 neighbourhood + 2) / 4, the last
    // column / row standing in for the missing one on odd sides.
    const uint32_t ow = (w + 1) / 2, oh = (h + 1) / 2;
    for (uint32_t oy = 0; oy < oh; oy++) {
        const size_t y0 = 2 * oy, y1 = std::min<size_t>(y0 + 1, h - 1);
        for (uint32_t ox = 0; ox < ow; ox++) {
            const size_t x0 = 2 * ox, x1 = std::min<size_t>(x0 + 1, w - 1);
            const uint32_t sum = (uint32_t)px[y0 * w + x0] + px[y0 * w + x1] + px[y1 * w + x0] +
                                 px[y1 * w + x1];
            out[(size_t)oy * ow + ox] = (uint8_t)((sum + 2) / 4);
        }
    }
    return MARS_HIP_OK;
}

void mars_free(void *p) { free(p); }

} // extern "C"

namespace {

void to_c_header(const Header &h, mars_header *o)
{
    *o = mars_header{};
    o->min_size = h.min_size;
    o->max_size = h.max_size;
    o->shift = h.shift;
    o->bits_alfa = h.bits_alfa;
    o->bits_beta = h.bits_beta;
    o->int_max_alfa = h.int_max_alfa;
    o->width = h.width;
    o->height = h.height;
    o->residual_qstep_bits = h.qstep_bits;
}

Header from_c_header(const mars_header &c)
{
    Header h{};
    h.min_size = c.min_size;
    h.max_size = c.max_size;
    h.shift = c.shift;
    h.bits_alfa = c.bits_alfa;
    h.bits_beta = c.bits_beta;
    h.int_max_alfa = c.int_max_alfa;
    h.width = c.width;
    h.height = c.height;
    h.qstep_bits = c.residual_qstep_bits;
    return h;
}

// Leaves (and their residual levels) into malloc'd C arrays.
int export_leaves(const std::vector<Leaf> &leaves, mars_leaf **out, size_t *n_out,
                  int32_t **residual, size_t *n_residual)
{
    *out = nullptr;
    *n_out = 0;
    if (residual) *residual = nullptr;
    if (n_residual) *n_residual = 0;
    size_t nres = 0;
    for (const Leaf &l : leaves) nres += l.residual.size();
    mars_leaf *L = leaves.empty() ? nullptr : (mars_leaf *)malloc(sizeof(mars_leaf) * leaves.size());
    int32_t *R = nres && residual ? (int32_t *)malloc(sizeof(int32_t) * nres) : nullptr;
    if ((!leaves.empty() && !L) || (nres && residual && !R)) {
        free(L);
        free(R);
        return MARS_HIP_E_ARG;
    }
    size_t off = 0;
    for (size_t i = 0; i < leaves.size(); i++) {
        const Leaf &l = leaves[i];
        mars_leaf &c = L[i];
        c = mars_leaf{};
        c.row = l.row;
        c.col = l.col;
        c.size = l.size;
        c.mode = l.mode;
        c.qalfa = l.qalfa;
        c.qbeta = l.qbeta;
        c.isometry = l.isometry;
        c.dom_row = l.dom_row;
        c.dom_col = l.dom_col;
        c.qgx = l.qgx;
        c.qgy = l.qgy;
        c.residual_off = (uint32_t)off;
        c.residual_len = (uint32_t)l.residual.size();
        if (R && !l.residual.empty()) memcpy(R + off, l.residual.data(), 4 * l.residual.size());
        off += l.residual.size();
    }
    *out = L;
    *n_out = leaves.size();
    if (residual) *residual = R;
    if (n_residual) *n_residual = R ? nres : 0;
    return MARS_HIP_OK;
}

// C leaves back to Leaf, checked against the plane (decode_leaf indexes the
// previous image through every field checked here).
int import_leaves(const Header &h, const mars_leaf *in, size_t n, const int32_t *residual,
                  size_t n_residual, bool check_plane, std::vector<Leaf> &out)
{
    out.clear();
    out.reserve(n);
    for (size_t i = 0; i < n; i++) {
        const mars_leaf &c = in[i];
        Leaf l{};
        l.row = c.row;
        l.col = c.col;
        l.size = c.size;
        l.mode = (uint8_t)c.mode;
        l.qalfa = c.qalfa;
        l.qbeta = c.qbeta;
        l.isometry = (uint8_t)c.isometry;
        l.dom_row = c.dom_row;
        l.dom_col = c.dom_col;
        l.qgx = c.qgx;
        l.qgy = c.qgy;
        if (c.mode > 3 || c.isometry > 7 || c.size == 0) return MARS_HIP_E_ARG;
        if (c.residual_len) {
            if ((uint64_t)c.residual_off + c.residual_len > n_residual || !residual) return MARS_HIP_E_ARG;
            l.residual.assign(residual + c.residual_off, residual + c.residual_off + c.residual_len);
        }
        if (check_plane) {
            if ((uint64_t)c.row + c.size > h.height || (uint64_t)c.col + c.size > h.width)
                return MARS_HIP_E_ARG;
            if (c.qalfa != 0 && c.mode != 1 &&
                ((uint64_t)c.dom_row + 2ull * c.size > h.height ||
                 (uint64_t)c.dom_col + 2ull * c.size > h.width))
                return MARS_HIP_E_ARG;
            if (c.mode == 3 && c.residual_len != 0 && (uint64_t)c.residual_len != (uint64_t)c.size * c.size)
                return MARS_HIP_E_ARG;
        }
        out.push_back(std::move(l));
    }
    return MARS_HIP_OK;
}

} // namespace

extern "C" {

int mars_read(const uint8_t *data, size_t len, mars_header *hdr, mars_leaf **leaves,
              size_t *n_leaves, int32_t **residual, size_t *n_residual)
{
    if ((!data && len) || !hdr || !leaves || !n_leaves) return MARS_HIP_E_ARG;
    *leaves = nullptr;
    *n_leaves = 0;
    if (residual) *residual = nullptr;
    if (n_residual) *n_residual = 0;
    Header h{};
    std::vector<Leaf> ls;
    const int rc = format_read(data, len, h, ls);
    if (rc) return rc;
    to_c_header(h, hdr);
    return export_leaves(ls, leaves, n_leaves, residual, n_residual);
}

int mars_write(const mars_header *hdr, const mars_leaf *leaves, size_t n_leaves,
               const int32_t *residual, size_t n_residual, uint8_t **out, size_t *out_len)
{
    if (!hdr || (!leaves && n_leaves) || !out || !out_len) return MARS_HIP_E_ARG;
    *out = nullptr;
    *out_len = 0;
    const Header h = from_c_header(*hdr);
    std::vector<Leaf> ls;
    int rc = import_leaves(h, leaves, n_leaves, residual, n_residual, false, ls);
    if (rc) return rc;
    std::vector<uint8_t> bytes;
    rc = format_write(h, ls, bytes);
    if (rc) return rc;
    uint8_t *buf = (uint8_t *)malloc(std::max<size_t>(bytes.size(), 1));
    if (!buf) return MARS_HIP_E_ARG;
    memcpy(buf, bytes.data(), bytes.size());
    *out = buf;
    *out_len = bytes.size();
    return MARS_HIP_OK;
}

int mars_ifs_parse(const uint8_t *data, size_t len, mars_header *hdr, mars_leaf **leaves,
                   size_t *n_leaves)
{
    if ((!data && len) || !hdr || !leaves || !n_leaves) return MARS_HIP_E_ARG;
    *leaves = nullptr;
    *n_leaves = 0;
    Header h{};
    std::vector<Leaf> ls;
    const int rc = ifs_parse(data, len, h, ls);
    if (rc) return rc;
    to_c_header(h, hdr);
    return export_leaves(ls, leaves, n_leaves, nullptr, nullptr);
}

int mars_decode_iterative(const mars_header *hdr, const mars_leaf *leaves, size_t n_leaves,
                          const int32_t *residual, size_t n_residual, uint32_t iterations,
                          uint8_t *out)
{
    if (!hdr || (!leaves && n_leaves) || !out) return MARS_HIP_E_ARG;
    const Header h = from_c_header(*hdr);
    if (h.width == 0 || h.height == 0 || h.bits_alfa > 31 || h.bits_beta > 31)
        return MARS_HIP_E_ARG;
    std::vector<Leaf> ls;
    const int rc = import_leaves(h, leaves, n_leaves, residual, n_residual, true, ls);
    if (rc) return rc;
    const size_t n = (size_t)h.width * h.height;
    std::vector<uint8_t> img(n, 128), next(n);
    for (uint32_t it = 0; it < iterations; it++) {
        std::fill(next.begin(), next.end(), 0);
        for (const Leaf &l : ls) decode_leaf(h, l, img.data(), h.width, next.data());
        img.swap(next);
    }
    memcpy(out, img.data(), n);
    return MARS_HIP_OK;
}

} // extern "C"
