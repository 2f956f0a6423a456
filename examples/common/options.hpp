// Command line and JSON configuration of the example programs: the behaviour
// of the reference's examples/common/cli.{hpp,cpp} without Boost.
//
//   program INPUT OUTPUT [--name value ...] [--config settings.json]
//
// Long options only ("-h" is --help); "--name value" or "--name=value"; a
// multi-value option takes every following token up to the next "--" option,
// so negative numbers are values. The first two bare tokens are the input and
// output paths. --config reads a JSON object whose keys are long option names
// (strings, numbers, booleans or arrays of them); an option given on the
// command line wins over the file. Unknown keys, "config" and "help" in the
// file are errors, like cli.cpp:62-67.
#pragma once

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <iterator>
#include <map>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cli {

struct option {
	option(std::string aName, std::string aHelp, std::optional<std::string> aDefault = std::nullopt,
		bool aRequired = false, bool aMultitoken = false, bool aFlag = false)
		: name(std::move(aName)), help(std::move(aHelp)), default_value(std::move(aDefault)), required(aRequired),
		  multitoken(aMultitoken), flag(aFlag) {}
	std::string name;
	std::string help;
	std::optional<std::string> default_value;
	bool required;
	bool multitoken;
	bool flag;
};

// ---------------------------------------------------------------- JSON ----
namespace json {

// A parsed JSON value; numbers keep their source text (the option parser
// converts them like command-line tokens).
struct value {
	enum kind_t { null_v, boolean, number, string, array, object } kind = null_v;
	std::string text;                                   // boolean ("true"/"false"), number, string
	std::vector<value> items;                           // array
	std::vector<std::pair<std::string, value>> members; // object, in file order
};

class parser {
public:
	explicit parser(const std::string& aText) : mText(aText) {}

	value parse_document()
	{
		value v = parse_value();
		skip_space();
		if (mPos != mText.size()) {
			fail("unexpected data after the JSON value");
		}
		return v;
	}

private:
	[[noreturn]] void fail(const std::string& aWhat) const
	{
		throw std::invalid_argument(aWhat + " at offset " + std::to_string(mPos));
	}
	void skip_space()
	{
		while (mPos < mText.size() && std::isspace(static_cast<unsigned char>(mText[mPos]))) {
			++mPos;
		}
	}
	bool consume(char aChar)
	{
		skip_space();
		if (mPos < mText.size() && mText[mPos] == aChar) {
			++mPos;
			return true;
		}
		return false;
	}
	void expect(char aChar)
	{
		if (!consume(aChar)) {
			fail(std::string("expected '") + aChar + "'");
		}
	}
	void append_utf8(std::string& aOut, unsigned aCode)
	{
		if (aCode < 0x80) {
			aOut += static_cast<char>(aCode);
		} else if (aCode < 0x800) {
			aOut += static_cast<char>(0xC0 | (aCode >> 6));
			aOut += static_cast<char>(0x80 | (aCode & 0x3F));
		} else if (aCode < 0x10000) {
			aOut += static_cast<char>(0xE0 | (aCode >> 12));
			aOut += static_cast<char>(0x80 | ((aCode >> 6) & 0x3F));
			aOut += static_cast<char>(0x80 | (aCode & 0x3F));
		} else {
			aOut += static_cast<char>(0xF0 | (aCode >> 18));
			aOut += static_cast<char>(0x80 | ((aCode >> 12) & 0x3F));
			aOut += static_cast<char>(0x80 | ((aCode >> 6) & 0x3F));
			aOut += static_cast<char>(0x80 | (aCode & 0x3F));
		}
	}
	unsigned hex4()
	{
		if (mPos + 4 > mText.size()) {
			fail("truncated \\u escape");
		}
		unsigned code = 0;
		for (int i = 0; i < 4; ++i) {
			const char c = mText[mPos++];
			code <<= 4;
			if (c >= '0' && c <= '9') {
				code |= static_cast<unsigned>(c - '0');
			} else if (c >= 'a' && c <= 'f') {
				code |= static_cast<unsigned>(c - 'a' + 10);
			} else if (c >= 'A' && c <= 'F') {
				code |= static_cast<unsigned>(c - 'A' + 10);
			} else {
				fail("invalid \\u escape");
			}
		}
		return code;
	}
	std::string parse_string()
	{
		expect('"');
		std::string out;
		while (true) {
			if (mPos >= mText.size()) {
				fail("unterminated string");
			}
			const char c = mText[mPos++];
			if (c == '"') {
				return out;
			}
			if (static_cast<unsigned char>(c) < 0x20) {
				fail("control character in string");
			}
			if (c != '\\') {
				out += c;
				continue;
			}
			if (mPos >= mText.size()) {
				fail("unterminated escape");
			}
			const char e = mText[mPos++];
			switch (e) {
			case '"': out += '"'; break;
			case '\\': out += '\\'; break;
			case '/': out += '/'; break;
			case 'b': out += '\b'; break;
			case 'f': out += '\f'; break;
			case 'n': out += '\n'; break;
			case 'r': out += '\r'; break;
			case 't': out += '\t'; break;
			case 'u': {
				unsigned code = hex4();
				if (code >= 0xD800 && code < 0xDC00 && mPos + 1 < mText.size() && mText[mPos] == '\\' &&
					mText[mPos + 1] == 'u') {
					mPos += 2;
					const unsigned low = hex4();
					code = 0x10000 + ((code - 0xD800) << 10) + (low - 0xDC00);
				}
				append_utf8(out, code);
				break;
			}
			default: fail("invalid escape");
			}
		}
	}
	value parse_number()
	{
		const std::size_t start = mPos;
		if (mPos < mText.size() && mText[mPos] == '-') {
			++mPos;
		}
		const auto digits = [&] {
			const std::size_t from = mPos;
			while (mPos < mText.size() && std::isdigit(static_cast<unsigned char>(mText[mPos]))) {
				++mPos;
			}
			return mPos > from;
		};
		if (!digits()) {
			fail("invalid number");
		}
		if (mPos < mText.size() && mText[mPos] == '.') {
			++mPos;
			if (!digits()) {
				fail("invalid number");
			}
		}
		if (mPos < mText.size() && (mText[mPos] == 'e' || mText[mPos] == 'E')) {
			++mPos;
			if (mPos < mText.size() && (mText[mPos] == '+' || mText[mPos] == '-')) {
				++mPos;
			}
			if (!digits()) {
				fail("invalid number");
			}
		}
		value v;
		v.kind = value::number;
		v.text = mText.substr(start, mPos - start);
		return v;
	}
	bool literal(const char* aWord)
	{
		const std::string word(aWord);
		if (mText.compare(mPos, word.size(), word) == 0) {
			mPos += word.size();
			return true;
		}
		return false;
	}
	value parse_value()
	{
		skip_space();
		if (mPos >= mText.size()) {
			fail("expected a value");
		}
		value v;
		const char c = mText[mPos];
		if (c == '{') {
			++mPos;
			v.kind = value::object;
			if (consume('}')) {
				return v;
			}
			do {
				skip_space();
				std::string key = parse_string();
				expect(':');
				v.members.emplace_back(std::move(key), parse_value());
			} while (consume(','));
			expect('}');
		} else if (c == '[') {
			++mPos;
			v.kind = value::array;
			if (consume(']')) {
				return v;
			}
			do {
				v.items.push_back(parse_value());
			} while (consume(','));
			expect(']');
		} else if (c == '"') {
			v.kind = value::string;
			v.text = parse_string();
		} else if (literal("true")) {
			v.kind = value::boolean;
			v.text = "true";
		} else if (literal("false")) {
			v.kind = value::boolean;
			v.text = "false";
		} else if (literal("null")) {
			v.kind = value::null_v;
		} else {
			v = parse_number();
		}
		return v;
	}

	const std::string& mText;
	std::size_t mPos = 0;
};

inline value parse(const std::string& aText) { return parser(aText).parse_document(); }

} // namespace json

// ------------------------------------------------------------- options ----
class arguments {
public:
	explicit arguments(std::vector<option> aOptions) : mOptions(std::move(aOptions)) {}

	const std::vector<option>& options() const { return mOptions; }
	bool has(const std::string& aName) const { return mValues.count(aName) > 0; }
	const std::vector<std::string>& tokens(const std::string& aName) const
	{
		const auto found = mValues.find(aName);
		if (found == mValues.end()) {
			throw std::invalid_argument("the option '--" + aName + "' has no value");
		}
		return found->second.tokens;
	}
	std::string text(const std::string& aName) const { return tokens(aName).front(); }
	int integer(const std::string& aName) const { return to_int(text(aName), aName); }
	float real(const std::string& aName) const
	{
		const auto& token = text(aName);
		char* end = nullptr;
		const float v = std::strtof(token.c_str(), &end);
		if (token.empty() || end != token.c_str() + token.size()) {
			throw std::invalid_argument("the argument ('" + token + "') for option '--" + aName + "' is invalid");
		}
		return v;
	}
	std::vector<int> integers(const std::string& aName) const
	{
		std::vector<int> out;
		for (const auto& token : tokens(aName)) {
			out.push_back(to_int(token, aName));
		}
		return out;
	}

	/// Parse the command line; returns false when --help was given (help printed).
	bool parse(int aArgc, char** aArgv, const std::string& aUsage, int aPositionals)
	{
		std::vector<std::string> positional;
		bool options_done = false;
		for (int i = 1; i < aArgc; ++i) {
			std::string token = aArgv[i];
			if (!options_done && token == "--") {
				options_done = true;
				continue;
			}
			if (!options_done && token == "-h") {
				token = "--help";
			}
			if (options_done || token.size() < 3 || token.compare(0, 2, "--") != 0) {
				positional.push_back(token);
				continue;
			}
			std::string name = token.substr(2);
			std::optional<std::string> inline_value;
			if (const auto eq = name.find('='); eq != std::string::npos) {
				inline_value = name.substr(eq + 1);
				name = name.substr(0, eq);
			}
			const option& spec = find(name, "unrecognised option '" + token + "'");
			if (mValues.count(name)) {
				throw std::invalid_argument("option '--" + name + "' cannot be specified more than once");
			}
			entry e;
			if (spec.flag) {
				e.tokens.push_back("true");
			} else if (inline_value) {
				e.tokens.push_back(*inline_value);
			} else if (spec.multitoken) {
				while (i + 1 < aArgc && std::string(aArgv[i + 1]).compare(0, 2, "--") != 0) {
					e.tokens.push_back(aArgv[++i]);
				}
			} else if (i + 1 < aArgc) {
				e.tokens.push_back(aArgv[++i]);
			}
			if (e.tokens.empty()) {
				throw std::invalid_argument("the required argument for option '--" + name + "' is missing");
			}
			mValues[name] = std::move(e);
		}
		if (has("help")) {
			print_help(aUsage);
			return false;
		}
		const char* names[] = {"input", "output"};
		const int first = aPositionals == 2 ? 0 : 1;
		if (static_cast<int>(positional.size()) > aPositionals) {
			throw std::invalid_argument("too many positional options have been specified on the command line");
		}
		for (std::size_t k = 0; k < positional.size(); ++k) {
			const std::string name = names[first + k];
			if (mValues.count(name)) {
				throw std::invalid_argument("option '--" + name + "' cannot be specified more than once");
			}
			mValues[name] = entry{{positional[k]}, false};
		}
		if (has("config")) {
			read_config(text("config"));
		}
		for (const auto& spec : mOptions) {
			if (!mValues.count(spec.name) && spec.default_value) {
				mValues[spec.name] = entry{{*spec.default_value}, true};
			}
			if (spec.required && !mValues.count(spec.name)) {
				throw std::invalid_argument("the option '--" + spec.name + "' is required but missing");
			}
		}
		return true;
	}

private:
	struct entry {
		std::vector<std::string> tokens;
		bool from_file = false;
	};

	const option& find(const std::string& aName, const std::string& aError) const
	{
		for (const auto& spec : mOptions) {
			if (spec.name == aName) {
				return spec;
			}
		}
		throw std::invalid_argument(aError);
	}

	static int to_int(const std::string& aToken, const std::string& aName)
	{
		char* end = nullptr;
		const long v = std::strtol(aToken.c_str(), &end, 10);
		if (aToken.empty() || end != aToken.c_str() + aToken.size() || v < -2147483647L - 1 || v > 2147483647L) {
			throw std::invalid_argument("the argument ('" + aToken + "') for option '--" + aName + "' is invalid");
		}
		return static_cast<int>(v);
	}

	static std::string token_of(const json::value& aValue)
	{
		if (aValue.kind == json::value::string || aValue.kind == json::value::number ||
			aValue.kind == json::value::boolean) {
			return aValue.text;
		}
		throw std::invalid_argument("Expected a string, number or boolean");
	}

	void read_config(const std::string& aPath)
	{
		std::ifstream file(aPath);
		if (!file) {
			throw std::runtime_error(aPath + ": unable to open configuration");
		}
		try {
			const std::string contents{std::istreambuf_iterator<char>(file), {}};
			const auto document = json::parse(contents);
			if (document.kind != json::value::object) {
				throw std::invalid_argument("Expected a JSON object");
			}
			for (const auto& [key, item] : document.members) {
				const option* spec = nullptr;
				for (const auto& o : mOptions) {
					if (o.name == key) {
						spec = &o;
					}
				}
				if (!spec || key == "config" || key == "help") {
					throw std::invalid_argument("Unknown configuration option '" + key + "'");
				}
				if (mValues.count(key)) {
					continue; // the command line wins
				}
				entry e;
				e.from_file = true;
				try {
					if (item.kind == json::value::array) {
						for (const auto& element : item.items) {
							e.tokens.push_back(token_of(element));
						}
					} else {
						e.tokens.push_back(token_of(item));
					}
					if (e.tokens.empty() || (!spec->multitoken && e.tokens.size() != 1)) {
						throw std::invalid_argument("Wrong number of values");
					}
				} catch (const std::exception& error) {
					throw std::invalid_argument("Option '" + key + "': " + error.what());
				}
				mValues[key] = std::move(e);
			}
		} catch (const std::exception& error) {
			throw std::runtime_error(aPath + ": " + error.what());
		}
	}

	void print_help(const std::string& aUsage) const
	{
		std::cout << aUsage << ":\n";
		for (const auto& spec : mOptions) {
			std::string left = "  --" + spec.name + (spec.flag ? "" : spec.multitoken ? " arg ..." : " arg");
			if (spec.default_value) {
				left += " (=" + *spec.default_value + ")";
			}
			if (left.size() < 34) {
				left.resize(34, ' ');
			} else {
				left += ' ';
			}
			std::cout << left << spec.help << '\n';
		}
	}

	std::vector<option> mOptions;
	std::map<std::string, entry> mValues;
};

} // namespace cli
