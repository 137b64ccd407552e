// The reference's own MathTest / PatternTest cases, run against the C++
// drop-in layer (cassandra_amd/cpp/include, the reference's API) -- plus fit
// checks that need the GPU.  `dropin_test host` runs the cases that need no
// device, `dropin_test gpu` the rest.  Exit status 0 when every case passed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "IO/Binary.h"
#include "Math/Math.h"
#include "Pattern/Evaluator.h"
#include "Pattern/Helpers.h"
#include "Pattern/Indexer.h"
#include "Pattern/Names.h"
#include "PatternFit/PatternFit.h"

static int failures = 0;
#define CHECK(cond)                                                                 \
    do {                                                                            \
        if (!(cond)) {                                                              \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);         \
            failures++;                                                             \
            return;                                                                 \
        }                                                                           \
    } while (0)

struct Case {
    const char* name;
    bool gpu;
    std::function<void()> fn;
};

static std::vector<Case>& cases()
{
    static std::vector<Case> c;
    return c;
}
#define TEST_CASE(name, gpu) \
    static void name();      \
    static bool reg_##name = (cases().push_back({#name, gpu, name}), true); \
    static void name()

// ---- PatternTest/Pattern.h:4-107 test masks and symmetric variants
static const std::vector<uint64_t> test_masks{0x3, 0x8040, 0x102, 0x81, 0x0100000000000001ULL,
                                              0x4200000000000042ULL, 0x4080000000000102ULL, 0x8100000000000081ULL};

static std::vector<Position> variants(Position p)
{
    return {p,
            FlippedCodiagonal(p),
            FlippedDiagonal(p),
            FlippedHorizontal(p),
            FlippedVertical(p),
            FlippedCodiagonal(FlippedHorizontal(p)),
            FlippedDiagonal(FlippedHorizontal(p)),
            FlippedVertical(FlippedHorizontal(p))};
}

static std::vector<Position> random_positions(unsigned seed, int n)
{
    std::mt19937_64 eng(seed);
    std::vector<Position> out;
    for (int i = 0; i < n; i++) {
        uint64_t a = eng(), b = eng();
        out.emplace_back(a & ~b, ~a & b);
    }
    return out;
}

// ---- MathTest/Vector.cpp
TEST_CASE(vector_known_answers, false)
{
    CHECK((Vector(3, 1.0f) + Vector(3, 2.0f)) == Vector(3, 3.0f));
    CHECK((Vector(3, 2.0f) - Vector(3, 1.0f)) == Vector(3, 1.0f));
    CHECK((Vector(3, 2.0f) * 3.0f) == Vector(3, 6.0f));
    CHECK((Vector(3, 6.0f) / 2.0f) == Vector(3, 3.0f));
    CHECK(elementwise_multiplication(Vector(3, 2.0f), Vector(3, 3.0f)) == Vector(3, 6.0f));
    CHECK(elementwise_division(Vector(3, 6.0f), Vector(3, 2.0f)) == Vector(3, 3.0f));
    CHECK(inv(Vector{1.0f, 2.0f, 4.0f}) == (Vector{1.0f, 0.5f, 0.25f}));
    CHECK(dot(Vector(3, 2.0f), Vector(3, 3.0f)) == 18.0f);
    CHECK(L1_norm(Vector(3, -2.0f)) == 6.0f);
    CHECK(L2_norm(Vector{3.0f, 4.0f}) == 5.0f);
    CHECK(sum(Vector(3, 2.0f)) == 6.0f);
    bool threw = false;
    try {
        Vector v(3, 2.0f);
        v += Vector(4, 3.0f);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw);
}

// ---- BoardTest/BitBoard.cpp:8-23
TEST_CASE(flip_known_answers, false)
{
    CHECK(FlippedCodiagonal(0xFULL) == 0x8080808000000000ULL);
    CHECK(FlippedDiagonal(0xFULL) == 0x01010101ULL);
    CHECK(FlippedHorizontal(0xFULL) == 0xF0ULL);
    CHECK(FlippedVertical(0xFULL) == 0x0F00000000000000ULL);
    CHECK(PossibleMoves(Position::Start()) == 0x0000102004080000ULL);
}

// ---- PatternTest/Indexer.cpp:8-47
TEST_CASE(indexer_symmetry_independence, false)
{
    auto check = [](auto& ix) {
        for (Position pos : random_positions(1337, 10000)) {
            auto base = ix.Indices(pos);
            std::sort(base.begin(), base.end());
            for (Position v : variants(pos)) {
                auto vi = ix.Indices(v);
                std::sort(vi.begin(), vi.end());
                if (vi != base) return false;
            }
        }
        return true;
    };
    for (uint64_t m : test_masks) {
        auto ix = CreateIndexer(m);
        CHECK(check(*ix));
    }
    GroupIndexer g(test_masks);
    CHECK(check(g));
}

// ---- PatternTest/Indexer.cpp:51-92
TEST_CASE(indexer_covers_index_space, false)
{
    for (uint64_t m : test_masks) {
        auto ix = CreateIndexer(m);
        std::set<uint32_t> seen;
        for (Position c : Configurations(m))
            for (uint32_t i : ix->Indices(c)) seen.insert(i);
        for (int i = 0; i < ix->index_space_size; i++) CHECK(seen.count(uint32_t(i)) == 1);
    }
    CHECK(ConfigurationCount(pattern::edax) == 114363);
    GroupIndexer g(pattern::edax);
    CHECK(g.index_space_size == 114363 && g.Variations().size() == 46);
}

// ---- PatternTest/Evaluator.cpp:7-31
TEST_CASE(score_estimator_round_trip_and_symmetry, false)
{
    for (uint64_t m : test_masks) {
        std::vector<float> w(ConfigurationCount(m));
        for (std::size_t i = 0; i < w.size(); i++) w[i] = float(i + 1);
        ScoreEstimator est({m}, w);
        CHECK(est.Weights() == w);
        for (Position c : Configurations(m)) {
            const float base = est.Eval(c);
            for (Position v : variants(c)) CHECK(est.Eval(v) == base);
        }
    }
}

// ---- IO/Binary.cpp:3-68: the weight file round trip and its byte layout
TEST_CASE(model_file_round_trip, false)
{
    const std::vector<uint64_t> pat{pattern::L02X, pattern::D8};
    PatternBasedEstimator est(20, pat);
    CHECK(est.Stages() == 4);
    const std::size_t m = ConfigurationCount(pat);
    for (int s = 0; s < est.Stages(); s++) {
        std::vector<float> w(m);
        for (std::size_t i = 0; i < m; i++) w[i] = float(s) + 0.25f * float(i % 97);
        est.score.estimators[s] = ScoreEstimator(pat, w);
    }
    const auto file = std::filesystem::temp_directory_path() / "cad_model_round_trip.bin";
    Save(est, file);
    PatternBasedEstimator back = LoadPatternBasedEstimator(file);
    CHECK(back.StageSize() == 20 && back.Stages() == 4 && back.Pattern() == est.Pattern());
    for (int s = 0; s < 4; s++) CHECK(back.score.estimators[s].Weights() == est.score.estimators[s].Weights());
    CHECK(back.accuracy.ParameterValues() == est.accuracy.ParameterValues());
    // i32 stage size, then vectors as u64 count + elements
    const std::size_t n = est.Pattern().size();
    const std::size_t want = 4 + (8 + 8 * n) + 8 + 4 * ((8 + 8 * n) + (8 + 4 * m)) + (8 + 4 * 5);
    CHECK(std::filesystem::file_size(file) == want);
    std::ifstream in(file, std::ios::binary);
    int32_t stage = 0;
    uint64_t count = 0;
    in.read(reinterpret_cast<char*>(&stage), 4);
    in.read(reinterpret_cast<char*>(&count), 8);
    CHECK(stage == 20 && count == n);
    std::filesystem::remove(file);
}

// ---- MathTest/MatrixCSR.cpp:3-54 (products on the GPU)
static MatrixCSR kat_matrix()
{
    MatrixCSR m(2, 3, 4);
    m.Row(0)[0] = 0;
    m.Row(0)[1] = 1;
    m.Row(1)[0] = 1;
    m.Row(1)[1] = 2;
    m.Row(2)[0] = 3;
    m.Row(2)[1] = 3;
    return m;
}

TEST_CASE(matrix_csr_known_answers, true)
{
    MatrixCSR m = kat_matrix();
    CHECK(m.Rows() == 3 && m.Cols() == 4 && m.Row(0).size() == 2);
    CHECK((m * Vector{1, 2, 3, 4}) == (Vector{3, 5, 8}));
    CHECK((Vector{1, 2, 3} * m) == (Vector{1, 3, 2, 6}));
    CHECK(JacobiPreconditionerOfATA(m) == (Vector{1, 0.5f, 1, 0.25f}));
    // rows changed after a product: the device copy follows
    m.Row(2)[1] = 0;
    CHECK((m * Vector{1, 2, 3, 4}) == (Vector{3, 5, 5}));
    bool threw = false;
    try {
        m* Vector(3, 1.0f);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw);
}

static std::vector<ScoredPosition> scored(int n, uint64_t seed)
{
    std::mt19937_64 eng(seed);
    std::vector<ScoredPosition> out;
    // random-play like positions: use the start position plus random legal moves
    for (int i = 0; i < n; i++) {
        Position p = Position::Start();
        for (int ply = 0; ply < 30; ply++) {
            uint64_t mv = PossibleMoves(p);
            if (!mv) {
                Position q(p.Opponent(), p.Player());
                if (!PossibleMoves(q)) break;
                p = q;
                continue;
            }
            int r = int(eng() % uint64_t(std::popcount(mv)));
            for (int j = 0; j < r; j++) mv &= mv - 1;
            const int m = std::countr_zero(mv);
            const uint64_t f = Flips(p, m);
            p = Position(p.Opponent() ^ f, p.Player() ^ f ^ (1ULL << m));
        }
        out.push_back({p, Score(int(eng() % 65) - 32)});
    }
    return out;
}

// FitWeights on the GPU against the reference-structured PCG of Solver.h run
// on host vectors (each product on the GPU): same algorithm, same float type.
TEST_CASE(fit_weights_against_host_pcg, true)
{
    const std::vector<uint64_t> pat{pattern::L02X, pattern::Q3, pattern::D8};
    auto sp = scored(20000, 7);
    Vector w = FitWeights(pat, sp, 10);
    GroupIndexer g(pat);
    MatrixCSR A(g.Variations().size(), sp.size(), std::size_t(g.index_space_size));
    for (std::size_t i = 0; i < sp.size(); i++) g.InsertIndices(sp[i].pos, A.Row(i));
    Vector b(sp.size());
    for (std::size_t i = 0; i < sp.size(); i++) b[i] = float(int8_t(sp[i].score));
    DiagonalPreconditioner P(JacobiPreconditionerOfATA(A));
    PCG solver(transposed(A) * A, P, Vector(A.Cols(), 0.0f), transposed(A) * b);
    solver.Iterate(10);
    // The float dots of the host template move its iterate by about 1e-4..1e-3
    // (relative) after 10 steps, so both are measured against the FP64 PCG
    // iterate, and FitWeights must be no farther from it than the host PCG.
    std::vector<double> b64(b.begin(), b.end());
    DeviceSolveOptions o;
    o.method = DeviceSolveOptions::PCG;
    o.precision = DeviceSolveOptions::FP64;
    o.max_iterations = 10;
    o.tol = 0.0;
    o.check_every = 10;
    const auto x64 = SolveOnDevice(A, b64, o);
    Vector xe(x64.begin(), x64.end());
    const Vector pg = A * w, ph = A * solver.X(), pe = A * xe;
    double dg = 0, dh = 0, den = 0;
    for (std::size_t i = 0; i < pg.size(); i++) {
        dg += double(pg[i] - pe[i]) * (pg[i] - pe[i]);
        dh += double(ph[i] - pe[i]) * (ph[i] - pe[i]);
        den += double(pe[i]) * pe[i];
    }
    CHECK(std::sqrt(dh / den) < 1e-2);
    CHECK(std::sqrt(dg) <= std::max(2.0 * std::sqrt(dh), 2e-4 * std::sqrt(den)));
}

TEST_CASE(eval_batch_matches_eval, true)
{
    const std::vector<uint64_t> pat{pattern::L02X, pattern::B5, pattern::Q3, pattern::D8};
    std::vector<float> w(ConfigurationCount(pat));
    std::mt19937 eng(3);
    std::normal_distribution<float> nd;
    for (auto& v : w) v = nd(eng);
    ScoreEstimator est(pat, w);
    auto sp = scored(3000, 11);
    std::vector<Position> pos;
    for (auto& s : sp) pos.push_back(s.pos);
    const auto batch = est.EvalBatch(pos);
    for (std::size_t i = 0; i < pos.size(); i++) CHECK(batch[i] == est.Eval(pos[i]));
}

TEST_CASE(device_lsqr_reaches_tolerance, true)
{
    const std::vector<uint64_t> pat{pattern::L02X, pattern::B5, pattern::Q3, pattern::D8};
    auto sp = scored(30000, 5);
    GroupIndexer g(pat);
    MatrixCSR A(g.Variations().size(), sp.size(), std::size_t(g.index_space_size));
    for (std::size_t i = 0; i < sp.size(); i++) g.InsertIndices(sp[i].pos, A.Row(i));
    std::vector<double> b(sp.size());
    for (std::size_t i = 0; i < sp.size(); i++) b[i] = double(int8_t(sp[i].score));
    DeviceSolveReport rep;
    auto x = SolveOnDevice(A, b, {}, &rep);
    CHECK(rep.converged && rep.rel_normal_residual <= 1e-10 && x.size() == A.Cols());
    // the FitStages loop over empty-count stages
    PatternBasedEstimator est(10, pat);
    FitStages(est, sp, 5);
    CHECK(est.Stages() == 7);
}

// FitStages fits all stages in one batch; each stage must get the weights
// FitWeights gives its positions alone (PatternFit/PatternFit.cpp:72-81).
TEST_CASE(fit_stages_equals_stagewise_fit_weights, true)
{
    const std::vector<uint64_t> pat{pattern::L02X, pattern::Q3, pattern::D8};
    auto sp = scored(12000, 9);
    // spread the empty counts: play on from some of the positions
    for (std::size_t i = 0; i < sp.size(); i += 3) {
        Position p = sp[i].pos;
        for (int ply = 0; ply < int(i % 17); ply++) {
            uint64_t mv = PossibleMoves(p);
            if (!mv) break;
            const int m = std::countr_zero(mv);
            const uint64_t f = Flips(p, m);
            p = Position(p.Opponent() ^ f, p.Player() ^ f ^ (1ULL << m));
        }
        sp[i].pos = p;
    }
    PatternBasedEstimator est(5, pat);
    FitStages(est, sp, 8);
    int checked = 0;
    for (int stage = 0; stage < est.Stages(); stage++) {
        std::vector<ScoredPosition> part;
        for (auto& s : sp)
            if (s.EmptyCount() / 5 == stage) part.push_back(s);
        if (part.size() < 200) continue;
        const Vector w = FitWeights(pat, part, 8);
        ScoreEstimator single(pat, w);
        double num = 0, den = 0;
        for (auto& s : part) {
            const double a = est.score.estimators[stage].Eval(s.pos), b = single.Eval(s.pos);
            num += (a - b) * (a - b);
            den += b * b;
        }
        CHECK(std::sqrt(num / den) < 1e-4);
        checked++;
    }
    CHECK(checked >= 2);
}

// ---- file modes for the oracle-parity tests in Python ----------------------
// All little-endian: u64 counts, then the elements.
template <class T>
static T get(std::ifstream& f)
{
    T v{};
    f.read(reinterpret_cast<char*>(&v), sizeof(T));
    if (!f) throw std::runtime_error("truncated input");
    return v;
}
template <class T>
static std::vector<T> get_run(std::ifstream& f)
{
    std::vector<T> v(get<uint64_t>(f));
    f.read(reinterpret_cast<char*>(v.data()), std::streamsize(sizeof(T) * v.size()));
    if (!f) throw std::runtime_error("truncated input");
    return v;
}
template <class T>
static void put_run(std::ofstream& f, const std::vector<T>& v)
{
    const uint64_t n = v.size();
    f.write(reinterpret_cast<const char*>(&n), sizeof n);
    f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(sizeof(T) * v.size()));
}

// fit-stages: in = i32 stage size, i32 iterations, run of u64 masks, run of
// u64 {P, O} pairs, run of i8 scores; out = u64 stages, then per stage the
// run of f32 compact weights FitStages gave it.
static int fit_stages_file(const char* in_path, const char* out_path)
{
    std::ifstream in(in_path, std::ios::binary);
    const int stage_size = get<int32_t>(in), iterations = get<int32_t>(in);
    const auto masks = get_run<uint64_t>(in);
    const auto po = get_run<uint64_t>(in);
    const auto score = get_run<int8_t>(in);
    if (po.size() != 2 * score.size()) throw std::runtime_error("positions and scores differ in count");
    std::vector<ScoredPosition> sp(score.size());
    for (std::size_t i = 0; i < sp.size(); i++) sp[i] = ScoredPosition{Position(po[2 * i], po[2 * i + 1]), score[i]};
    PatternBasedEstimator est(stage_size, masks);
    FitStages(est, sp, iterations);
    std::ofstream out(out_path, std::ios::binary);
    const uint64_t n = uint64_t(est.Stages());
    out.write(reinterpret_cast<const char*>(&n), sizeof n);
    for (int s = 0; s < est.Stages(); s++) put_run(out, est.score.estimators[s].Weights());
    return out ? 0 : 1;
}

// eval-batch: in = run of u64 masks, run of f32 compact weights, run of u64
// {P, O} pairs; out = run of f32 ScoreEstimator::EvalBatch values.
static int eval_batch_file(const char* in_path, const char* out_path)
{
    std::ifstream in(in_path, std::ios::binary);
    const auto masks = get_run<uint64_t>(in);
    const auto w = get_run<float>(in);
    const auto po = get_run<uint64_t>(in);
    std::vector<Position> pos(po.size() / 2);
    for (std::size_t i = 0; i < pos.size(); i++) pos[i] = Position(po[2 * i], po[2 * i + 1]);
    ScoreEstimator est(masks, w);
    const std::vector<float> v = est.EvalBatch(pos);
    std::ofstream out(out_path, std::ios::binary);
    put_run(out, v);
    return out ? 0 : 1;
}

int main(int argc, char** argv)
{
    const std::string mode = argc > 1 ? argv[1] : "host";
    if ((mode == "fit-stages" || mode == "eval-batch") && argc == 4) {
        try {
            return mode == "fit-stages" ? fit_stages_file(argv[2], argv[3]) : eval_batch_file(argv[2], argv[3]);
        } catch (const std::exception& e) {
            std::printf("%s failed: %s\n", mode.c_str(), e.what());
            return 1;
        }
    }
    if (mode == "copy-model" && argc == 4) {  // load a weight file and save it again
        try {
            Save(LoadPatternBasedEstimator(argv[2]), argv[3]);
        } catch (const std::exception& e) {
            std::printf("copy-model failed: %s\n", e.what());
            return 1;
        }
        return 0;
    }
    int run = 0;
    for (auto& c : cases()) {
        if (c.gpu != (mode == "gpu")) continue;
        const int before = failures;
        try {
            c.fn();
        } catch (const std::exception& e) {
            std::printf("  FAILED %s: exception %s\n", c.name, e.what());
            failures++;
        }
        std::printf("%s %s\n", failures == before ? "ok  " : "FAIL", c.name);
        std::fflush(stdout);
        run++;
    }
    std::printf("%d cases, %d failures\n", run, failures);
    return failures ? 1 : 0;
}
