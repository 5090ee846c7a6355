"""Mouse::make_source and Mouse::update (src/mouse.cu:4-45) without SDL:
closed-form checks of the Source a frame gets from the pointer state."""
from opensayal_amd import Mouse


def test_idle_mouse_is_inactive_with_starting_speed():
    src = Mouse().make_source(1080, 1)
    assert src.active == 0 and src.velocity == 200.0 and src.smoke == 0.0


def test_left_button_paints_smoke_at_the_cell_under_the_pointer():
    m = Mouse()
    m.motion(37, 10)
    m.button_down(1)
    src = m.make_source(100, 3)
    # 37 / 3 = 12, 10 / 3 = 3 (C division), y flipped: 100 - 3
    assert (src.active, src.smoke, src.velocity) == (1, 1.0, 200.0)
    assert (src.position_x, src.position_y) == (12, 97)


def test_middle_button_reverses_the_wheel_speed():
    m = Mouse()
    m.button_down(2)
    m.wheel(3)
    assert m.make_source(50, 1).velocity == -230.0
    m.button_up()
    assert m.make_source(50, 1).active == 0


def test_wheel_never_goes_below_zero():
    m = Mouse()
    m.wheel(-20)  # 200 - 200 = 0 is allowed
    assert m.wheel_value == 0
    m.wheel(-1)   # -10 is refused
    assert m.wheel_value == 0
    m.wheel(2)
    assert m.wheel_value == 20


def test_negative_positions_truncate_toward_zero():
    m = Mouse()
    m.motion(-5, -7)
    src = m.make_source(40, 2)
    assert (src.position_x, src.position_y) == (-2, 43)
