#include "seg_tuning.h"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

#include "hip_check.h"

namespace mgsort {

namespace {

// A JSON value: just enough of the format for the table (objects, arrays,
// strings without escapes other than \" and \\, numbers, literals).
struct Json {
  enum Kind { kNull, kBool, kNum, kStr, kArr, kObj } kind = kNull;
  double num = 0;
  bool b = false;
  std::string str;
  std::vector<Json> arr;
  std::map<std::string, Json> obj;
  const Json* get(const std::string& k) const {
    auto it = obj.find(k);
    return it == obj.end() ? nullptr : &it->second;
  }
};

class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s) {}
  Json Parse() {
    Json v = Value();
    Ws();
    if (i_ != s_.size()) Fail("trailing characters");
    return v;
  }

 private:
  [[noreturn]] void Fail(const char* what) {
    throw Error(std::string("tuning table: ") + what + " at byte " + std::to_string(i_));
  }
  void Ws() {
    while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\n' || s_[i_] == '\r' || s_[i_] == '\t')) ++i_;
  }
  bool Eat(char c) {
    Ws();
    if (i_ < s_.size() && s_[i_] == c) {
      ++i_;
      return true;
    }
    return false;
  }
  std::string Str() {
    if (!Eat('"')) Fail("expected a string");
    std::string out;
    while (i_ < s_.size() && s_[i_] != '"') {
      if (s_[i_] == '\\') {
        if (++i_ >= s_.size()) Fail("bad escape");
        if (s_[i_] != '"' && s_[i_] != '\\' && s_[i_] != '/') Fail("unsupported escape");
      }
      out.push_back(s_[i_++]);
    }
    if (i_ >= s_.size()) Fail("unterminated string");
    ++i_;
    return out;
  }
  Json Value() {
    Ws();
    if (i_ >= s_.size()) Fail("unexpected end");
    Json v;
    const char c = s_[i_];
    if (c == '{') {
      ++i_;
      v.kind = Json::kObj;
      if (Eat('}')) return v;
      do {
        Ws();
        std::string k = Str();
        if (!Eat(':')) Fail("expected ':'");
        v.obj[k] = Value();
      } while (Eat(','));
      if (!Eat('}')) Fail("expected '}'");
    } else if (c == '[') {
      ++i_;
      v.kind = Json::kArr;
      if (Eat(']')) return v;
      do v.arr.push_back(Value());
      while (Eat(','));
      if (!Eat(']')) Fail("expected ']'");
    } else if (c == '"') {
      v.kind = Json::kStr;
      v.str = Str();
    } else if (s_.compare(i_, 4, "true") == 0 || s_.compare(i_, 5, "false") == 0) {
      v.kind = Json::kBool;
      v.b = s_[i_] == 't';
      i_ += v.b ? 4 : 5;
    } else if (s_.compare(i_, 4, "null") == 0) {
      i_ += 4;
    } else {
      const char* p = s_.c_str() + i_;
      char* end = nullptr;
      v.num = std::strtod(p, &end);
      if (end == p) Fail("unexpected character");
      v.kind = Json::kNum;
      i_ += (size_t)(end - p);
    }
    return v;
  }
  const std::string& s_;
  size_t i_ = 0;
};

int Int(const Json& o, const char* k) {
  const Json* v = o.get(k);
  if (v == nullptr || v->kind != Json::kNum) throw Error(std::string("tuning table: entry without numeric ") + k);
  return (int)v->num;
}

std::string Text(const Json& o, const char* k) {
  const Json* v = o.get(k);
  if (v == nullptr || v->kind != Json::kStr) throw Error(std::string("tuning table: entry without string ") + k);
  return v->str;
}

}  // namespace

SegTuningTable SegTuningTable::Load(const std::string& path) {
  SegTuningTable t;
  if (path.empty()) return t;
  std::ifstream f(path);
  if (!f) throw Error("tuning table: cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  const Json root = Parser(text).Parse();
  const Json* fmt = root.get("format");
  const Json* entries = root.get("entries");
  if (fmt == nullptr || fmt->kind != Json::kNum || fmt->num != 1 || entries == nullptr || entries->kind != Json::kArr)
    throw Error("tuning table: " + path + " is not a format-1 segmented-sort table");
  static const std::map<std::string, mgs_key> keys = {{"u32", MGS_KEY_U32}, {"u64", MGS_KEY_U64},
                                                      {"i32", MGS_KEY_I32}, {"i64", MGS_KEY_I64},
                                                      {"f32", MGS_KEY_F32}, {"f64", MGS_KEY_F64}};
  static const std::map<std::string, mgs_val> vals = {{"none", MGS_VAL_NONE}, {"u32", MGS_VAL_U32},
                                                      {"u64", MGS_VAL_U64}};
  for (const Json& e : entries->arr) {
    const auto k = keys.find(Text(e, "key"));
    const auto v = vals.find(Text(e, "val"));
    const Json* tn = e.get("tuning");
    if (k == keys.end() || v == vals.end() || tn == nullptr || tn->kind != Json::kObj)
      throw Error("tuning table: bad entry in " + path);
    SegTuningEntry x{};
    x.kt = k->second;
    x.vt = v->second;
    x.size_log2 = Int(e, "size_log2");
    x.tuning.large_variant = Int(*tn, "large_variant");
    x.tuning.small_max = Int(*tn, "small_max");
    x.tuning.medium_max = Int(*tn, "medium_max");
    x.tuning.max_segment_keys = tn->get("max_segment_keys") ? Int(*tn, "max_segment_keys") : 0;
    if (x.tuning.large_variant < 0 || x.tuning.large_variant >= mgs_segmented_variant_count())
      throw Error("tuning table: variant out of range in " + path);
    t.entries_.push_back(x);
  }
  return t;
}

const mgs_tuning* SegTuningTable::Lookup(mgs_key kt, mgs_val vt, size_t total_keys) const {
  const int bin = (int)std::lround(std::log2((double)(total_keys ? total_keys : 1)));
  const SegTuningEntry* best = nullptr;
  for (const SegTuningEntry& e : entries_) {
    if (e.kt != kt || e.vt != vt) continue;
    if (best == nullptr || std::abs(e.size_log2 - bin) < std::abs(best->size_log2 - bin) ||
        (std::abs(e.size_log2 - bin) == std::abs(best->size_log2 - bin) && e.size_log2 < best->size_log2))
      best = &e;
  }
  return best ? &best->tuning : nullptr;
}

}  // namespace mgsort
