// config.h — WHIR protocol configuration derived from the protocol parameters.
//
// Restates whir::Config::from_params (protocols/whir/whir.hpp:118-224) and
// irs_commit::Config::from_params (protocols/irs_commit.hpp:267-324) with the
// same double-precision operations in the same order, so the in/out-domain
// sample counts and PoW thresholds come out identical.  The target field is the
// Goldilocks cubic extension (192 field bits, goldilocks_ext3.hpp:37).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace whir_host {

struct IrsConfig {
    size_t num_vectors = 0, vector_size = 0, codeword_length = 0, depth = 0;
    size_t in_domain = 0, out_domain = 0;
    double johnson_slack = 0.0;

    size_t num_cols() const { return num_vectors * depth; }
    size_t message_length() const { return vector_size / depth; }
    double rate() const { return (double)message_length() / (double)codeword_length; }
    bool unique_decoding() const { return out_domain == 0 && johnson_slack == 0.0; }
    double list_size() const {
        if (unique_decoding()) return 1.0;
        return 1.0 / (2.0 * johnson_slack * std::sqrt(rate()));
    }
    double rbr_queries() const {
        double per_query_agreement = unique_decoding() ? ((1.0 + rate()) * 0.5) : (std::sqrt(rate()) + johnson_slack);
        return (double)in_domain * (-std::log2(per_query_agreement));
    }
    double rbr_fold_prox_gaps(double fbits) const {
        double log_inv_rate_bits = -std::log2(rate());
        double log_message_length = std::log2((double)message_length());
        double error_bits;
        if (unique_decoding()) {
            error_bits = log_message_length + log_inv_rate_bits;
        } else {
            const double LOG2_10 = 3.321928094887362;
            error_bits = 7.0 * LOG2_10 + 3.5 * log_inv_rate_bits + 2.0 * log_message_length;
        }
        return fbits - error_bits;
    }

    static IrsConfig derive(double security, bool unique_dec, size_t num_vec, size_t vec_size, size_t depth,
                            double rate_val, double fbits) {
        IrsConfig c;
        c.num_vectors = num_vec;
        c.vector_size = vec_size;
        c.depth = depth;
        size_t message_len = vec_size / depth;
        c.codeword_length = (size_t)std::ceil((double)message_len / rate_val);
        double actual_rate = (double)message_len / (double)c.codeword_length;
        c.johnson_slack = unique_dec ? 0.0 : std::sqrt(actual_rate) / 20.0;
        if (unique_dec) {
            c.out_domain = 0;
        } else {
            double L = 1.0 / (2.0 * c.johnson_slack * std::sqrt(actual_rate));
            double list_pairs = L * (L - 1.0) / 2.0;
            double bits_per_sample = fbits - std::log2((double)(vec_size - 1));
            c.out_domain = (size_t)std::ceil((security + std::log2(list_pairs)) / bits_per_sample);
            if (c.out_domain < 1) c.out_domain = 1;
        }
        double per_query_agreement = unique_dec ? ((1.0 + actual_rate) * 0.5) : (std::sqrt(actual_rate) + c.johnson_slack);
        c.in_domain = (size_t)std::ceil(security / (-std::log2(per_query_agreement)));
        return c;
    }
};

struct SumcheckConfig {
    size_t initial_size = 0, num_rounds = 0;
    uint64_t round_pow = UINT64_MAX;  // threshold; UINT64_MAX = no PoW
    size_t final_size() const { return initial_size >> num_rounds; }
};

struct RoundConfig {
    IrsConfig irs;
    SumcheckConfig sumcheck;
    uint64_t pow = UINT64_MAX;
};

struct Params {
    bool unique_decoding = false;
    size_t starting_log_inv_rate = 1, initial_folding_factor = 4, folding_factor = 4;
    size_t security_level = 128, pow_bits = 20, batch_size = 1;
};

struct WhirConfig {
    IrsConfig initial;
    SumcheckConfig initial_sumcheck;
    uint64_t initial_skip_pow = UINT64_MAX;
    std::vector<RoundConfig> rounds;
    SumcheckConfig final_sumcheck;
    uint64_t final_pow = UINT64_MAX;

    // PoW threshold for a difficulty in bits: ceil(2^(64 - d)), none for d <= 0
    static uint64_t pow_threshold(double d) {
        if (d <= 0.0) return UINT64_MAX;
        double t = std::ceil(std::exp2(64.0 - d));
        if (t >= 18446744073709551616.0) return UINT64_MAX;
        return (uint64_t)t;
    }

    static WhirConfig derive(size_t size, const Params& p) {
        const double fbits = 192.0;
        WhirConfig c;
        double security = (double)p.security_level;
        double protocol_security = (double)(p.security_level - p.pow_bits);
        size_t log_inv_rate = p.starting_log_inv_rate;
        size_t num_vars = 0;
        for (size_t s = size; s > 1; s >>= 1) ++num_vars;
        double initial_rate = std::pow(0.5, (double)p.starting_log_inv_rate);
        c.initial = IrsConfig::derive(protocol_security, p.unique_decoding, p.batch_size, size,
                                      size_t{1} << p.initial_folding_factor, initial_rate, fbits);
        double start_fold_bits;
        {
            double prox_gap_bits = c.initial.rbr_fold_prox_gaps(fbits);
            double list_bits = std::log2(c.initial.list_size());
            start_fold_bits = std::max(security - std::min(prox_gap_bits, fbits - list_bits - 1.0), 0.0);
        }
        double skip_bits;
        {
            double prox_gap_bits = c.initial.rbr_fold_prox_gaps(fbits) + std::log2((double)p.initial_folding_factor);
            skip_bits = std::max(security - prox_gap_bits, 0.0);
        }
        c.initial_sumcheck.initial_size = size;
        c.initial_sumcheck.num_rounds = p.initial_folding_factor;
        c.initial_sumcheck.round_pow = pow_threshold(start_fold_bits);
        c.initial_skip_pow = pow_threshold(skip_bits);

        size_t round_index = 0;
        size_t prev_in_domain = c.initial.in_domain;
        double prev_query_error = c.initial.rbr_queries();
        num_vars -= p.initial_folding_factor;
        while (num_vars >= p.folding_factor) {
            size_t round_fold = round_index == 0 ? p.initial_folding_factor : p.folding_factor;
            size_t next_log_inv_rate = log_inv_rate + (round_fold - 1);
            double round_rate = std::pow(0.5, (double)next_log_inv_rate);
            RoundConfig round;
            round.irs = IrsConfig::derive(protocol_security, p.unique_decoding, 1, size_t{1} << num_vars,
                                          size_t{1} << p.folding_factor, round_rate, fbits);
            double combination_bits;
            {
                double list_bits = std::log2(round.irs.list_size());
                size_t constraint_count = round.irs.out_domain + prev_in_domain;
                combination_bits = fbits - (std::log2((double)constraint_count) + list_bits + 1.0);
            }
            double round_pow_bits = std::max(security - std::min(prev_query_error, combination_bits), 0.0);
            double fold_pow_bits;
            {
                double prox_gap_bits = round.irs.rbr_fold_prox_gaps(fbits);
                double list_bits = std::log2(round.irs.list_size());
                fold_pow_bits = std::max(security - std::min(prox_gap_bits, fbits - list_bits - 1.0), 0.0);
            }
            round.sumcheck.initial_size = size_t{1} << num_vars;
            round.sumcheck.num_rounds = p.folding_factor;
            round.sumcheck.round_pow = pow_threshold(fold_pow_bits);
            round.pow = pow_threshold(round_pow_bits);
            ++round_index;
            num_vars -= p.folding_factor;
            log_inv_rate = next_log_inv_rate;
            prev_in_domain = round.irs.in_domain;
            prev_query_error = round.irs.rbr_queries();
            c.rounds.push_back(round);
        }
        double last_query_error = c.rounds.empty() ? c.initial.rbr_queries() : c.rounds.back().irs.rbr_queries();
        double final_pow_bits = std::max(security - last_query_error, 0.0);
        double final_fold_pow_bits = std::max(security - fbits + 1.0, 0.0);
        c.final_sumcheck.initial_size = size_t{1} << num_vars;
        c.final_sumcheck.num_rounds = num_vars;
        c.final_sumcheck.round_pow = pow_threshold(final_fold_pow_bits);
        c.final_pow = pow_threshold(final_pow_bits);
        return c;
    }

    // Flat form: [n_rounds | initial irs (7) | initial sumcheck size, rounds, pow |
    // skip pow | per_query_agreement round: irs (7), sumcheck size, rounds, pow, round pow |
    // final sumcheck size, rounds, pow | final pow]; irs = num_vectors,
    // vector_size, codeword_length, depth, in_domain, out_domain, slack bits.
    std::vector<uint64_t> flat() const {
        std::vector<uint64_t> o;
        auto irs = [&o](const IrsConfig& c) {
            uint64_t bits;
            std::memcpy(&bits, &c.johnson_slack, 8);
            for (uint64_t v : {(uint64_t)c.num_vectors, (uint64_t)c.vector_size, (uint64_t)c.codeword_length,
                               (uint64_t)c.depth, (uint64_t)c.in_domain, (uint64_t)c.out_domain, bits})
                o.push_back(v);
        };
        o.push_back(rounds.size());
        irs(initial);
        o.push_back(initial_sumcheck.initial_size);
        o.push_back(initial_sumcheck.num_rounds);
        o.push_back(initial_sumcheck.round_pow);
        o.push_back(initial_skip_pow);
        for (const auto& r : rounds) {
            irs(r.irs);
            o.push_back(r.sumcheck.initial_size);
            o.push_back(r.sumcheck.num_rounds);
            o.push_back(r.sumcheck.round_pow);
            o.push_back(r.pow);
        }
        o.push_back(final_sumcheck.initial_size);
        o.push_back(final_sumcheck.num_rounds);
        o.push_back(final_sumcheck.round_pow);
        o.push_back(final_pow);
        return o;
    }
};

}  // namespace whir_host
