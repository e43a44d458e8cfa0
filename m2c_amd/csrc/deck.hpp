// deck.hpp -- the parsed form of an M2C input deck, shared by the readers of its objects (host code).
// input_deck.cpp holds the parser (the grammar of parser/Parser.y) and the solver options; output.cpp
// reads the Output object from the same tree.
#pragma once
#include <initializer_list>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/m2c_hip.h"

namespace m2c_deck {

struct DeckError {
  int code;
  std::string msg;
};

// throws DeckError{code, "line <line>: <m>"} (no line prefix when line <= 0)
[[noreturn]] void fail(int code, int line, const std::string &m);

struct Value {
  enum Kind { INT, DBL, SYM, STR } kind = INT;
  long long i = 0;
  double d = 0.0;
  std::string s;
  int line = 0;
};

// An object of the deck: its fields, and its member objects (array elements under "#<index>").
struct Node {
  std::string type;   // name of the object (aliases resolved); "#" for an element of an array
  std::map<std::string, Value> leaves;
  std::map<std::string, std::unique_ptr<Node>> kids;
  int line = 0;       // first mention

  Node *kid(const std::string &n, int at) {
    auto &p = kids[n];
    if (!p) {
      p.reset(new Node());
      p->type = n[0] == '#' ? "#" : n;
      p->line = at;
    }
    return p.get();
  }
  const Node *find(const char *n) const {
    auto it = kids.find(n);
    return it == kids.end() ? nullptr : it->second.get();
  }
  const Node *elem(long long i) const { return find(("#" + std::to_string(i)).c_str()); }
  bool empty() const { return leaves.empty() && kids.empty(); }
};

// the whole deck (throws DeckError)
void parse(const char *text, Node &root);
// the solver options, run controls and checked shapes of a parsed deck; returns an m2c status and
// fills `report` with the error or with the ignored options (one per line)
int read_config(const char *text, Node &root, m2c_config &c, m2c_deck_run &r, std::string &report);

// field readers with the reference's assignment rules (all do nothing when the field is absent)
const Value *leaf(const Node *n, const char *name);
void getd(const Node *n, const char *name, double &out);     // a double or an integer
void geti(const Node *n, const char *name, int &out);        // an integer
void gets(const Node *n, const char *name, std::string &out);  // a string
void gett(const Node *n, const char *name, int &out, std::initializer_list<std::pair<const char *, int>> opts,
          std::initializer_list<const char *> unsupported = {});   // a keyword
// elements of an array object in increasing index order
std::vector<const Node *> elements(const Node *arr);

// cell centres x[d] and widths w[d] of the three axes of a configuration (uniform, or stretched by
// its control points: m2c_mesh_1d); an m2c status
int mesh_axes(const m2c_config &c, std::vector<double> x[3], std::vector<double> w[3]);

}  // namespace m2c_deck
